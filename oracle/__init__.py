"""CPU oracle for the SCFT hot path -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``liboracle_scft.so`` (built from ``psx_oracle.cpp`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker.  The engine (``polyswift_plusplus_amd``) never imports it.

Each function names the reference lines it restates; the restatement itself
is in ``psx_oracle.cpp``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_scft.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


ORC_MAX_JOIN = 8
ORC_MAX_POLY = 8


class OrcBlock(ctypes.Structure):
    _fields_ = [
        ("species", ctypes.c_int), ("frac", ctypes.c_double), ("b", ctypes.c_double),
        ("nhead", ctypes.c_int), ("ntail", ctypes.c_int),
        ("head", ctypes.c_int * ORC_MAX_JOIN), ("tail", ctypes.c_int * ORC_MAX_JOIN),
    ]


@dataclass
class Block:
    """One block of a block graph: species 0 = A, 1 = B; lengthfrac; bSegRatio;
    indices of the blocks joined at its head and tail (empty = freeEnd)."""
    species: int
    frac: float
    b: float = 1.0
    head: tuple = ()
    tail: tuple = ()

    def c(self) -> OrcBlock:
        ob = OrcBlock()
        ob.species, ob.frac, ob.b = self.species, self.frac, self.b
        ob.nhead, ob.ntail = len(self.head), len(self.tail)
        for i, j in enumerate(self.head):
            ob.head[i] = j
        for i, j in enumerate(self.tail):
            ob.tail[i] = j
        return ob


class OrcPolymer(ctypes.Structure):
    _fields_ = [("volfrac", ctypes.c_double), ("lratio", ctypes.c_double), ("nblocks", ctypes.c_int)]


@dataclass
class Polymer:
    """One polymer of a blend: volfrac, lengthRatio (length / the first
    polymer's length) and its blocks (Block, joined by indices within the
    polymer)."""
    volfrac: float
    lratio: float
    blocks: tuple

    def c(self) -> OrcPolymer:
        return OrcPolymer(self.volfrac, self.lratio, len(self.blocks))


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("ds", ctypes.c_double), ("fA", ctypes.c_double), ("chiN", ctypes.c_double),
        ("bA", ctypes.c_double), ("bB", ctypes.c_double),
        ("mix", ctypes.c_int),
        ("lambda0", ctypes.c_double), ("lambda1", ctypes.c_double),
        ("depth", ctypes.c_int), ("beta", ctypes.c_double),
        ("anderson_start", ctypes.c_double),
        ("f_order", ctypes.c_int),
        ("quad", ctypes.c_int),
        ("nblocks", ctypes.c_int),
        ("blocks", ctypes.POINTER(OrcBlock)),
        ("npoly", ctypes.c_int),
        ("polys", ctypes.POINTER(OrcPolymer)),
    ]


@dataclass
class Params:
    """One bulk AB diblock melt.  Lengths in Rg, s in [0, 1]."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    ds: float
    fA: float
    chiN: float
    bA: float = 1.0
    bB: float = 1.0
    mix: int = 0            # 0 steepest descent, 1 Anderson
    lambda0: float = 0.1
    lambda1: float = 0.0
    depth: int = 0
    beta: float = 0.1
    anderson_start: float = 1e300
    f_order: int = 0        # 0 reference (F after update), 1 consistent
    quad: int = 0           # 0 Simpson 0.33333 (reference), 1 trapezoid
    blocks: tuple = ()      # Block graph; empty = the diblock of fA, bA, bB
    polymers: tuple = ()    # Polymer blend (replaces blocks); empty = one polymer

    @property
    def N(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    def c(self) -> OrcParams:
        cp = OrcParams(self.nx, self.ny, self.nz, self.dx, self.dy, self.dz,
                       self.ds, self.fA, self.chiN, self.bA, self.bB, self.mix,
                       self.lambda0, self.lambda1, self.depth, self.beta,
                       self.anderson_start, self.f_order, self.quad, 0, None, 0, None)
        blocks = self.all_blocks()
        if blocks:
            arr = (OrcBlock * len(blocks))(*[b.c() for b in blocks])
            cp.nblocks = len(blocks)
            cp.blocks = ctypes.cast(arr, ctypes.POINTER(OrcBlock))
            cp._keep = arr      # the array lives as long as the struct
        if self.polymers:
            parr = (OrcPolymer * len(self.polymers))(*[q.c() for q in self.polymers])
            cp.npoly = len(self.polymers)
            cp.polys = ctypes.cast(parr, ctypes.POINTER(OrcPolymer))
            cp._keep_p = parr
        return cp

    def all_blocks(self):
        if self.polymers:
            return tuple(b for q in self.polymers for b in q.blocks)
        return tuple(self.blocks)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_block_steps.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_block_steps.restype = ctypes.c_int
        L.orc_kfac.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 5 + [_dp]
        L.orc_fft3.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, ctypes.c_int]
        L.orc_dft3_brute.argtypes = [ctypes.c_int] * 3 + [_dp] * 4 + [ctypes.c_int]
        L.orc_scaled_fft_pair.argtypes = [ctypes.c_int] * 3 + [_dp] * 3
        L.orc_solve_q.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [ctypes.c_int, _dp]
        L.orc_quadrature.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_double,
                                     _dp, _dp, ctypes.c_double, _dp]
        L.orc_diblock_densities.argtypes = [ctypes.POINTER(OrcParams)] + [_dp] * 5
        L.orc_graph_densities.argtypes = [ctypes.POINTER(OrcParams)] + [_dp] * 5
        L.orc_graph_slice_count.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_graph_slice_count.restype = ctypes.c_long
        L.orc_species_fracs.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp]
        L.orc_sd_update.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_long] + [_dp] * 5 + [ctypes.c_double]
        L.orc_sd_update.restype = ctypes.c_double
        L.orc_free_energy.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_long] + [_dp] * 4 + [ctypes.c_double]
        L.orc_free_energy.restype = ctypes.c_double
        L.orc_free_energy_poly.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_long] + [_dp] * 5
        L.orc_free_energy_poly.restype = ctypes.c_double
        L.orc_polymer_count.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_polymer_count.restype = ctypes.c_int
        L.orc_scft_run.argtypes = ([ctypes.POINTER(OrcParams)] + [_dp] * 3 +
                                   [ctypes.c_int, ctypes.c_double] + [_dp] * 5)
        L.orc_scft_run_timed.argtypes = ([ctypes.POINTER(OrcParams)] + [_dp] * 3 +
                                         [ctypes.c_int] + [_dp] * 2)
        L.orc_box_stress.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp, _dp]
        L.orc_time_steps.argtypes = [ctypes.c_int] * 4
        L.orc_time_steps.restype = ctypes.c_double
        L.orc_k2.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [_dp]
        L.orc_k2.restype = None
        for nm in ("orc_fft_forward_abs", "orc_fft_forward_re", "orc_fft_backward_re"):
            getattr(L, nm).argtypes = [ctypes.c_int] * 3 + [_dp] * 2
        for nm in ("orc_fft_convolve_re", "orc_fft_scaled_pair_im"):
            getattr(L, nm).argtypes = [ctypes.c_int] * 3 + [_dp] * 3
        L.orc_spec_filter.argtypes = [ctypes.c_int] * 3 + [_dp, ctypes.c_double, ctypes.c_double,
                                                           ctypes.POINTER(ctypes.c_int), ctypes.c_double]
        L.orc_poisson_update.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [_dp] * 2
        L.orc_rng_create.restype = ctypes.c_void_p
        L.orc_rng_create.argtypes = [ctypes.c_ulong]
        L.orc_rng_destroy.argtypes = [ctypes.c_void_p]
        L.orc_rng_destroy.restype = None
        L.orc_rng_raw32.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.POINTER(ctypes.c_uint)]
        L.orc_rng_raw32.restype = None
        for nm in ("orc_rng_uniform", "orc_rng_gaussian"):
            getattr(L, nm).argtypes = [ctypes.c_void_p, ctypes.c_long, _dp]
            getattr(L, nm).restype = None
        L.orc_scft_run_noise.argtypes = ([ctypes.POINTER(OrcParams)] + [_dp] * 3 + [ctypes.c_int] +
                                         [_dp] * 3 + [ctypes.c_double, ctypes.c_void_p])
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def block_steps(length_frac: float, ds: float) -> int:
    """psbase/PsBlockBase.cpp:252-326; raises on a non-integral count."""
    r = lib().orc_block_steps(length_frac, ds)
    if r < 0:
        raise ValueError(f"non-integral or too few block steps: {length_frac}/{ds}")
    return r


def kfac(shape, cell, ds, b=1.0) -> np.ndarray:
    """exp(-ds b^2 k^2), full grid (psbase/PsGridField.cpp:66-104)."""
    out = np.empty(shape, dtype=np.float64)
    lib().orc_kfac(*shape, *cell, ds, b, _p(out))
    return out


def fft3(a: np.ndarray, sign: int = -1) -> np.ndarray:
    re, im = _f64(a.real).copy(), _f64(np.imag(a)).copy()
    if lib().orc_fft3(*a.shape, _p(re), _p(im), sign):
        raise ValueError("bad shape")
    return re + 1j * im


def dft3_brute(a: np.ndarray, sign: int = -1) -> np.ndarray:
    re, im = _f64(a.real), _f64(np.imag(a))
    ore, oim = np.empty_like(re), np.empty_like(im)
    lib().orc_dft3_brute(*a.shape, _p(re), _p(im), _p(ore), _p(oim), sign)
    return ore + 1j * oim


def scaled_fft_pair(data: np.ndarray, kdata: np.ndarray) -> np.ndarray:
    """psfft/PsFFTW.cpp:196-224 (unnormalised, real part kept)."""
    d, k = _f64(data), _f64(kdata)
    out = np.empty_like(d)
    lib().orc_scaled_fft_pair(*d.shape, _p(d), _p(k), _p(out))
    return out


def solve_q(kf: np.ndarray, wfac: np.ndarray, q0: np.ndarray, steps: int) -> np.ndarray:
    """pspolymer/PsFlexPseudoSpec.cpp:244-293: slices[0..steps]."""
    kf, wfac, q0 = _f64(kf), _f64(wfac), _f64(q0)
    out = np.empty((steps + 1,) + kf.shape)
    lib().orc_solve_q(*kf.shape, _p(kf), _p(wfac), _p(q0), steps, _p(out))
    return out


def quadrature(qs: np.ndarray, qts: np.ndarray, ds: float, Q: float) -> np.ndarray:
    """pspolymer/PsBlock.cpp:256-380 (0.33333 Simpson + trapezoid tail)."""
    qs, qts = _f64(qs), _f64(qts)
    steps = qs.shape[0] - 1
    out = np.empty(qs.shape[1:])
    lib().orc_quadrature(out.size, steps, ds, _p(qs), _p(qts), Q, _p(out))
    return out


def diblock_densities(P: Params, wA, wB, want_slices: bool = False):
    """Densities, Q (and optionally all slices qA, qB, qtB, qtA)."""
    wA, wB = _f64(wA), _f64(wB)
    phiA, phiB = np.empty_like(wA), np.empty_like(wB)
    Q = ctypes.c_double()
    cp = P.c()
    sl = None
    slp = None
    if want_slices:
        sA, sB = block_steps(P.fA, P.ds), block_steps(1.0 - P.fA, P.ds)
        sl = np.empty((2 * (sA + 1) + 2 * (sB + 1),) + wA.shape)
        slp = _p(sl)
    rc = lib().orc_diblock_densities(ctypes.byref(cp), _p(wA), _p(wB), _p(phiA),
                                     _p(phiB), ctypes.cast(ctypes.byref(Q), _dp), slp)
    if rc:
        raise RuntimeError(f"oracle densities failed rc={rc}")
    if want_slices:
        sA, sB = block_steps(P.fA, P.ds), block_steps(1.0 - P.fA, P.ds)
        i0 = 0
        qA = sl[i0:i0 + sA + 1]; i0 += sA + 1
        qB = sl[i0:i0 + sB + 1]; i0 += sB + 1
        qtB = sl[i0:i0 + sB + 1]; i0 += sB + 1
        qtA = sl[i0:i0 + sA + 1]
        return phiA, phiB, Q.value, (qA, qB, qtB, qtA)
    return phiA, phiB, Q.value


def graph_densities(P: Params, wA, wB, want_slices: bool = False):
    """Densities and Q of the block graph (pspolymer/PsBlockCopolymer.cpp:189-313);
    with want_slices, also [(q from head, q from tail)] per block, each
    (steps+1, *shape) and indexed from its own starting end."""
    wA, wB = _f64(wA), _f64(wB)
    phiA, phiB = np.empty_like(wA), np.empty_like(wB)
    Qs = np.zeros(ORC_MAX_POLY)
    cp = P.c()
    total = lib().orc_graph_slice_count(ctypes.byref(cp))
    if total < 0:
        raise ValueError("bad block graph")
    sl = np.empty((total,) + wA.shape) if want_slices else None
    rc = lib().orc_graph_densities(ctypes.byref(cp), _p(wA), _p(wB), _p(phiA), _p(phiB), _p(Qs),
                                   _p(sl) if want_slices else None)
    if rc:
        raise RuntimeError(f"oracle graph densities failed rc={rc}")
    Q = Qs[:len(P.polymers)].copy() if P.polymers else float(Qs[0])
    if not want_slices:
        return phiA, phiB, Q
    out, i0 = [], 0
    if P.polymers:
        pairs = [(b, q.lratio) for q in P.polymers for b in q.blocks]
    else:
        pairs = [(b, 1.0) for b in (P.blocks or (Block(0, P.fA, P.bA, (), (1,)),
                                                  Block(1, 1.0 - P.fA, P.bB, (0,), ())))]
    for b, lr in pairs:
        M = block_steps(lr * b.frac, P.ds) + 1
        out.append((sl[i0:i0 + M], sl[i0 + M:i0 + 2 * M]))
        i0 += 2 * M
    return phiA, phiB, Q, out


def species_fracs(P: Params):
    """(f_A, f_B): lengthfrac summed over each species' blocks."""
    a, b = ctypes.c_double(), ctypes.c_double()
    cp = P.c()
    lib().orc_species_fracs(ctypes.byref(cp), ctypes.cast(ctypes.byref(a), _dp),
                            ctypes.cast(ctypes.byref(b), _dp))
    return a.value, b.value


def sd_update(P: Params, wA, wB, p, phiA, phiB, Q):
    """In-place steepest-descent update; returns F (reference order)."""
    cp = P.c()
    return lib().orc_sd_update(ctypes.byref(cp), wA.size, _p(wA), _p(wB), _p(p),
                               _p(_f64(phiA)), _p(_f64(phiB)), Q)


def free_energy(P: Params, wA, wB, phiA, phiB, Q) -> float:
    """F of fields and densities; Q is one value per polymer for a blend."""
    cp = P.c()
    if P.polymers:
        Qs = _f64(np.atleast_1d(Q))
        if Qs.size != len(P.polymers):
            raise ValueError("a blend needs one Q per polymer")
        return lib().orc_free_energy_poly(ctypes.byref(cp), wA.size, _p(_f64(wA)), _p(_f64(wB)),
                                          _p(_f64(phiA)), _p(_f64(phiB)), _p(Qs))
    return lib().orc_free_energy(ctypes.byref(cp), wA.size, _p(_f64(wA)), _p(_f64(wB)),
                                 _p(_f64(phiA)), _p(_f64(phiB)), Q)


def scft_run(P: Params, wA, wB, p=None, n_iters=1, tol=0.0):
    """Run iterations in place on copies; returns dict of results."""
    wA, wB = _f64(wA).copy(), _f64(wB).copy()
    p = np.zeros_like(wA) if p is None else _f64(p).copy()
    F = np.zeros(n_iters); Qh = np.zeros(n_iters); err = np.zeros(n_iters)
    phiA, phiB = np.empty_like(wA), np.empty_like(wA)
    cp = P.c()
    done = lib().orc_scft_run(ctypes.byref(cp), _p(wA), _p(wB), _p(p), n_iters, tol,
                              _p(F), _p(Qh), _p(err), _p(phiA), _p(phiB))
    if done < 0:
        raise RuntimeError("oracle scft_run failed")
    return dict(wA=wA, wB=wB, p=p, F=F[:done], Q=Qh[:done], err=err[:done],
                phiA=phiA, phiB=phiB, iters=done)


def scft_run_timed(P: Params, wA, wB, p=None, n_iters=1):
    """n_iters full iterations from (wA, wB, p), each timed on the host's
    steady clock: dict(wA, wB, p, F, seconds) with seconds per iteration."""
    wA, wB = _f64(wA).copy(), _f64(wB).copy()
    p = np.zeros_like(wA) if p is None else _f64(p).copy()
    F, t = np.zeros(n_iters), np.zeros(n_iters)
    cp = P.c()
    done = lib().orc_scft_run_timed(ctypes.byref(cp), _p(wA), _p(wB), _p(p), n_iters, _p(F), _p(t))
    if done != n_iters:
        raise RuntimeError("oracle scft_run_timed failed")
    return dict(wA=wA, wB=wB, p=p, F=F, seconds=t)


def box_stress(P: Params, wA, wB) -> np.ndarray:
    """dF/dL_d at fixed fields (no reference counterpart)."""
    out = np.zeros(3)
    cp = P.c()
    if lib().orc_box_stress(ctypes.byref(cp), _p(_f64(wA)), _p(_f64(wB)), _p(out)):
        raise RuntimeError("oracle stress failed")
    return out


def time_steps(shape, steps: int) -> float:
    return lib().orc_time_steps(*shape, steps)


# -- the rest of PsFFTBase and its consumers ----------------------------------
def k2(shape, cell) -> np.ndarray:
    """k^2 of PsGridField::calck2 (psbase/PsGridField.cpp:66-104)."""
    out = np.empty(shape, dtype=np.float64)
    lib().orc_k2(*shape, *cell, _p(out))
    return out


def _unary(name, data):
    d = _f64(data)
    out = np.empty_like(d)
    if getattr(lib(), name)(*d.shape, _p(d), _p(out)):
        raise ValueError("bad shape")
    return out


def forward_fft_abs(data) -> np.ndarray:
    """|FFT(data)|^2 (psfft/PsFFTW.cpp:121-146)."""
    return _unary("orc_fft_forward_abs", data)


def forward_fft_re(data) -> np.ndarray:
    """Re(FFT(data)) (psfft/PsFFTW.cpp:258-276)."""
    return _unary("orc_fft_forward_re", data)


def backward_fft_re(data) -> np.ndarray:
    """Re(IFFT(data + 0i)), unnormalised (psfft/PsFFTW.cpp:278-296)."""
    return _unary("orc_fft_backward_re", data)


def convolve_re(d1, d2) -> np.ndarray:
    """Re(IFFT(FFT(d1) FFT(d2))) (psfft/PsFFTW.cpp:148-188)."""
    a, b = _f64(d1), _f64(d2)
    out = np.empty_like(a)
    lib().orc_fft_convolve_re(*a.shape, _p(a), _p(b), _p(out))
    return out


def scaled_fft_pair_im(data, kdata) -> np.ndarray:
    """Re(IFFT(kdata FFT(0 + i data))) (psfft/PsFFTW.cpp:226-256)."""
    a, k = _f64(data), _f64(kdata)
    out = np.empty_like(a)
    lib().orc_fft_scaled_pair_im(*a.shape, _p(a), _p(k), _p(out))
    return out


def spec_filter(w, cutoff: float, strength: float, cells=None, vwall: float = 0.0) -> np.ndarray:
    """Filtered copy of w: PsSimpleSpecFilter (cells None) or PsMultiSpecFilter
    (pseffhamil/PsSimpleSpecFilter.cpp:74-144, PsMultiSpecFilter.cpp:123-267,
    PsSpecFilterUpdater.cpp:222-239)."""
    out = _f64(w).copy()
    c = None if cells is None else (ctypes.c_int * 3)(*[int(x) for x in cells])
    r = lib().orc_spec_filter(*out.shape, _p(out), float(cutoff), float(strength), c, float(vwall))
    if r:
        raise ValueError("bad shape or spectral cell sizes")
    return out


def poisson_update(charge, psi, lB: float, cell):
    """(shifted charge, new psi) of PsPoissonUpdater::update
    (pseffhamil/PsPoissonUpdater.cpp:171-289)."""
    c, p = _f64(charge).copy(), _f64(psi).copy()
    lib().orc_poisson_update(*c.shape, *[float(x) for x in cell], float(lB), _p(c), _p(p))
    return c, p


# -- the reference's random stream ----------------------------------------------
class Rng:
    """PsRandom (psstd/PsRandom.h:147-170,305-373; PsRandom.cpp:58-95):
    MT19937 seeded by init_genrand, uniform y / 2^32, Marsaglia polar
    Gaussians."""

    def __init__(self, seed: int):
        self._h = lib().orc_rng_create(int(seed))

    def raw32(self, n: int) -> np.ndarray:
        out = np.empty(int(n), dtype=np.uint32)
        lib().orc_rng_raw32(self._h, int(n), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
        return out

    def uniform(self, n: int) -> np.ndarray:
        out = np.empty(int(n))
        lib().orc_rng_uniform(self._h, int(n), _p(out))
        return out

    def gaussian(self, n: int) -> np.ndarray:
        out = np.empty(int(n))
        lib().orc_rng_gaussian(self._h, int(n), _p(out))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_rng_destroy(self._h)
            self._h = None


def scft_run_noise(P: Params, wA, wB, p=None, n_iters=1, strength=0.0, rng: Rng = None):
    """Steepest descent with the reference's noise term drawn from rng
    (pseffhamil/PsPolymerUpdater.cpp:264-295)."""
    wA, wB = _f64(wA).copy(), _f64(wB).copy()
    p = np.zeros_like(wA) if p is None else _f64(p).copy()
    F = np.zeros(n_iters); Qh = np.zeros(n_iters); err = np.zeros(n_iters)
    cp = P.c()
    done = lib().orc_scft_run_noise(ctypes.byref(cp), _p(wA), _p(wB), _p(p), n_iters, _p(F), _p(Qh),
                                    _p(err), float(strength), rng._h)
    if done < 0:
        raise RuntimeError("oracle scft_run_noise failed")
    return dict(wA=wA, wB=wB, p=p, F=F, Q=Qh, err=err, iters=done)
