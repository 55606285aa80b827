"""Host-side mirror of the reference operator interface for the SCFT hot path.

Names and argument meanings follow PolySwift++:

* ``PsFFT.scaledFFTPair(data, kdata)`` -- ``PsFFTBase::scaledFFTPair``
  (psbase/PsFFTBase.h:105-106, psfft/PsFFTW.cpp:196-224): unnormalised
  forward/backward pair, real part kept.
* ``DiblockSCFT`` -- the AB diblock of ``PsBlockCopolymer`` with two
  ``PsFlexPseudoSpec`` blocks (pspolymer/PsBlockCopolymer.cpp:189-239),
  the ``flory`` interaction and the ``steepestDescent`` updater
  (pseffhamil/PsSteepDUpdater.cpp:159-333), plus Anderson mixing.
* ``BlockCopolymerSCFT`` -- the same for any tree of A and B ``Block``s
  joined by ``headjoined`` / ``tailjoined`` (linear multiblocks, stars).

Every call goes through the C ABI of libpsx_scft.so on the current HIP
device; errors raise ``PsxError`` the way the reference throws
``TxDebugExcept``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import PSX_FP32, PSX_FP64, PsxError, check

_dp = ctypes.POINTER(ctypes.c_double)


def _prec(p) -> int:
    if p in (PSX_FP64, "fp64", "f64", "double", np.float64):
        return PSX_FP64
    if p in (PSX_FP32, "fp32", "f32", "float", np.float32):
        return PSX_FP32
    raise ValueError(f"unknown precision {p!r}")


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class PsFFT:
    """FFT pair plugin over the psx HIP line-FFT kernels (serial plan)."""

    def __init__(self, shape, prec="fp64", stream=None):
        self._L = _abi.load()
        self.shape = tuple(int(s) for s in shape)
        self.prec = _prec(prec)
        self.dtype = np.float64 if self.prec == PSX_FP64 else np.float32
        h = ctypes.c_void_p()
        check(self._L.psx_fft_plan_create(*self.shape, self.prec, stream, ctypes.byref(h)))
        self._h = h

    def getFFTSize(self) -> int:
        return int(self._L.psx_fft_size(self._h))

    def scaledFFTPair(self, data, kdata) -> np.ndarray:
        """Host arrays in, host array out (the reference's own calling form)."""
        d = np.ascontiguousarray(data, dtype=self.dtype)
        k = np.ascontiguousarray(kdata, dtype=self.dtype)
        if d.size != self.getFFTSize() or k.size != d.size:
            raise PsxError(2, "data/kdata size differs from the FFT size")
        out = np.empty_like(d)
        check(self._L.psx_fft_scaled_pair_host(self._h, d.ctypes.data, k.ctypes.data,
                                                out.ctypes.data))
        return out.reshape(self.shape)

    def scaledFFTPair_device(self, d_data: int, d_kdata: int, d_res: int) -> None:
        """Device pointers (e.g. torch ``tensor.data_ptr()``); asynchronous."""
        check(self._L.psx_fft_scaled_pair(self._h, d_data, d_kdata, d_res))

    # -- the rest of PsFFTBase, host arrays in and out through device buffers --
    def _call(self, fn, arrays, n_out=1, extra=()):
        """Upload `arrays`, run fn(plan, *inputs, *outputs, *extra) on the
        device, and return the outputs (and the inputs, which some calls
        change in place) as host arrays of the grid's shape."""
        n = self.getFFTSize()
        ins = [np.ascontiguousarray(a, dtype=self.dtype).reshape(-1) for a in arrays]
        if any(a.size != n for a in ins):
            raise PsxError(2, "array size differs from the FFT size")
        nbytes = n * np.dtype(self.dtype).itemsize
        bufs = []
        try:
            for _ in range(len(ins) + n_out):
                d = ctypes.c_void_p()
                check(self._L.psx_buffer_create(nbytes, ctypes.byref(d)))
                bufs.append(d)
            for d, a in zip(bufs, ins):
                check(self._L.psx_buffer_upload(d, a.ctypes.data, nbytes))
            check(fn(self._h, *bufs, *extra))
            check(self._L.psx_fft_synchronize(self._h))
            outs = []
            for d in bufs:
                o = np.empty(n, dtype=self.dtype)
                check(self._L.psx_buffer_download(o.ctypes.data, d, nbytes))
                outs.append(o.reshape(self.shape))
            return outs
        finally:
            for d in bufs:
                self._L.psx_buffer_destroy(d)

    def forwardFFTAbs(self, data) -> np.ndarray:
        """|FFT(data)|^2 (the squared modulus, as PsFFTW::forwardFFTAbs)."""
        return self._call(self._L.psx_fft_forward_abs, [data])[-1]

    def convolveRe(self, data1, data2) -> np.ndarray:
        """Re(IFFT(FFT(data1) FFT(data2))), unnormalised."""
        return self._call(self._L.psx_fft_convolve_re, [data1, data2])[-1]

    def scaledFFTPairIm(self, data, kdata) -> np.ndarray:
        """Re(IFFT(kdata FFT(0 + i data))), unnormalised."""
        return self._call(self._L.psx_fft_scaled_pair_im, [data, kdata])[-1]

    def calcForwardFFT(self, data) -> np.ndarray:
        """Re(FFT(data)), unnormalised."""
        return self._call(self._L.psx_fft_forward_re, [data])[-1]

    def calcBackwardFFT(self, data) -> np.ndarray:
        """Re(IFFT(data + 0i)), unnormalised."""
        return self._call(self._L.psx_fft_backward_re, [data])[-1]

    def specFilter(self, w, cutoffFactor: float, filterStrength: float, specCellSizes=None,
                   wallVolume: float = 0.0) -> np.ndarray:
        """The filtered field (PsSimpleSpecFilter, or PsMultiSpecFilter with
        specCellSizes): mean removed, spectral mask, FFT pair / N."""
        cells = None if specCellSizes is None else (ctypes.c_int * 3)(*[int(c) for c in specCellSizes])

        def fn(h, d):
            return self._L.psx_fft_spec_filter(h, d, float(cutoffFactor), float(filterStrength), cells,
                                               float(wallVolume))
        return self._call(fn, [w], n_out=0)[0]

    def poissonUpdate(self, charge, psi, bjerrumLen: float, cell):
        """(charge with its mean removed, new psi) of PsPoissonUpdater::update."""
        dx, dy, dz = (float(c) for c in cell)

        def fn(h, dc, dp):
            return self._L.psx_fft_poisson_update(h, dc, dp, float(bjerrumLen), dx, dy, dz)
        c, p = self._call(fn, [charge, psi], n_out=0)
        return c, p

    def close(self):
        if getattr(self, "_h", None):
            self._L.psx_fft_plan_destroy(self._h)
            self._h = None

    __del__ = close


class Rng:
    """The reference's host random stream (PsRandom: MT19937, init_genrand
    seeding, uniform y / 2^32, Marsaglia polar Gaussians)."""

    def __init__(self, seed: int):
        self._L = _abi.load()
        h = ctypes.c_void_p()
        check(self._L.psx_rng_create(int(seed), ctypes.byref(h)))
        self._h = h

    def uniform(self, n: int) -> np.ndarray:
        out = np.empty(int(n))
        check(self._L.psx_rng_uniform(self._h, int(n), _ptr(out)))
        return out

    def gaussian(self, n: int) -> np.ndarray:
        out = np.empty(int(n))
        check(self._L.psx_rng_gaussian(self._h, int(n), _ptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._L.psx_rng_destroy(self._h)
            self._h = None

    __del__ = close


@dataclass
class Mixing:
    """Field-update method (``steepestDescent`` [lambda0, lambda1] or Anderson)."""
    kind: str = "steepest"       # "steepest" | "anderson"
    lambda0: float = 0.1
    lambda1: float = 0.0
    depth: int = 10
    beta: float = 0.1
    anderson_start: float = 1e300
    f_order: int = 0             # 0 reference (F after update), 1 consistent

    def c(self) -> _abi.Mixing:
        return _abi.Mixing(0 if self.kind == "steepest" else 1, self.lambda0, self.lambda1,
                           int(self.depth), self.beta, self.anderson_start, int(self.f_order))


class DiblockSCFT:
    """Bulk AB diblock melt on an nx*ny*nz periodic grid (one HIP device)."""

    def __init__(self, shape, cell, fA, ds, chiN, bA=1.0, bB=1.0, prec="fp64", stream=None,
                 quadrature="reference"):
        self._L = _abi.load()
        self.shape = tuple(int(s) for s in shape)
        self.cell = tuple(float(c) for c in cell)
        self.N = int(np.prod(self.shape))
        self.fA, self.ds, self.chiN = float(fA), float(ds), float(chiN)
        g = _abi.Grid(*self.shape, *self.cell)
        c = _abi.Diblock(self.fA, self.ds, float(bA), float(bB), self.chiN)
        h = ctypes.c_void_p()
        check(self._L.psx_scft_create(ctypes.byref(g), ctypes.byref(c), _prec(prec), stream,
                                      ctypes.byref(h)))
        self._h = h
        self._steps()
        if quadrature != "reference":
            self.setQuadrature(quadrature)

    def _steps(self):
        sA, sB = ctypes.c_int(), ctypes.c_int()
        check(self._L.psx_scft_block_steps(self._h, ctypes.byref(sA), ctypes.byref(sB)))
        self.sA, self.sB = sA.value, sB.value
        self.Ns = self.sA + self.sB

    # -- fields -------------------------------------------------------------
    def setFields(self, wA, wB, p=None) -> None:
        a, b = _f64(wA).reshape(-1), _f64(wB).reshape(-1)
        if a.size != self.N or b.size != self.N:
            raise PsxError(2, "field size differs from the grid")
        pp = None if p is None else _f64(p).reshape(-1)
        check(self._L.psx_scft_set_fields_host(self._h, _ptr(a), _ptr(b),
                                               None if pp is None else _ptr(pp)))

    def getFields(self):
        a, b, p = (np.empty(self.N) for _ in range(3))
        check(self._L.psx_scft_get_fields_host(self._h, _ptr(a), _ptr(b), _ptr(p)))
        return a.reshape(self.shape), b.reshape(self.shape), p.reshape(self.shape)

    def setFieldsDevice(self, d_wA: int, d_wB: int, d_p: int | None = None) -> None:
        check(self._L.psx_scft_set_fields(self._h, d_wA, d_wB, d_p))

    def setPrecision(self, prec) -> None:
        check(self._L.psx_scft_set_precision(self._h, _prec(prec)))

    def setNoise(self, strength: float, rng: "Rng | None") -> None:
        """Steepest-descent noise: strength * N(0, 1) per cell and step from
        rng (the reference's theRng); rng None turns it off."""
        self._noise_rng = rng       # the engine keeps the stream's pointer
        check(self._L.psx_scft_set_noise(self._h, float(strength), rng._h if rng else None))

    def specFilter(self, fields: str, cutoffFactor: float, filterStrength: float,
                   specCellSizes=None) -> None:
        """Spectral filter updater on the conjugate fields in place: fields
        "A", "B" or "AB" (PsSimpleSpecFilter, or PsMultiSpecFilter with
        specCellSizes)."""
        mask = (1 if "A" in fields else 0) | (2 if "B" in fields else 0)
        cells = None if specCellSizes is None else (ctypes.c_int * 3)(*[int(c) for c in specCellSizes])
        check(self._L.psx_scft_spec_filter(self._h, mask, float(cutoffFactor), float(filterStrength), cells))

    # -- polymer: PsBlockCopolymer::update split into its reference steps ----
    def solveQ(self) -> None:
        """w-factors and all block propagators (reset + solveQ per block)."""
        check(self._L.psx_scft_propagate(self._h))

    def calcBigQ(self) -> float:
        q = ctypes.c_double()
        check(self._L.psx_scft_Q(self._h, ctypes.byref(q)))
        return q.value

    def densities(self):
        """setCalcQQTIntegral + setPhysFields -> (phiA, phiB) on the host."""
        check(self._L.psx_scft_densities(self._h, None, None))
        return self.getDensities()

    def getDensities(self):
        """Densities of the last densities()/iterate() call."""
        a, b = np.empty(self.N), np.empty(self.N)
        check(self._L.psx_scft_get_densities_host(self._h, _ptr(a), _ptr(b)))
        return a.reshape(self.shape), b.reshape(self.shape)

    def update(self):
        """PsBlockCopolymer::update: propagate, Q, densities."""
        self.solveQ()
        return self.densities()

    def slice(self, path, t: int) -> np.ndarray:
        """Slice t of propagation path `path` (0 forward sweep, 1 backward
        sweep of a chain; a bool picks the backward sweep)."""
        out = np.empty(self.N)
        check(self._L.psx_scft_get_slice_host(self._h, int(path), int(t), _ptr(out)))
        return out.reshape(self.shape)

    def blockSlice(self, block: int, from_tail: bool, s: int) -> np.ndarray:
        """Slice s of a block's propagator from its head or from its tail
        (the reference's qs / qts of that block)."""
        out = np.empty(self.N)
        check(self._L.psx_scft_get_block_slice_host(self._h, int(block), int(bool(from_tail)), int(s),
                                                    _ptr(out)))
        return out.reshape(self.shape)

    def graphInfo(self):
        """(blocks, propagation paths, rounds of concurrent paths)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self._L.psx_scft_graph_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    # -- effective Hamiltonian ----------------------------------------------
    def steepestDescent(self, lambda0: float, lambda1: float = 0.0) -> float:
        """PsSteepDUpdater::update from the current densities; returns F."""
        F = ctypes.c_double()
        check(self._L.psx_scft_update_steepest(self._h, lambda0, lambda1, ctypes.byref(F)))
        return F.value

    def getFreeE(self) -> float:
        F = ctypes.c_double()
        check(self._L.psx_scft_free_energy(self._h, ctypes.byref(F)))
        return F.value

    def iterate(self, n_iters: int, mixing: Mixing | None = None, tol: float = 0.0,
                use_graph: bool = True) -> dict:
        """Full SCFT iterations (PsDomain::update); returns histories."""
        mixing = mixing or Mixing()
        F, Q, err = np.zeros(n_iters), np.zeros(n_iters), np.zeros(n_iters)
        done = ctypes.c_int()
        m = mixing.c()
        check(self._L.psx_scft_iterate(self._h, ctypes.byref(m), int(n_iters), float(tol),
                                       int(bool(use_graph)), _ptr(F), _ptr(Q), _ptr(err),
                                       ctypes.byref(done)))
        k = done.value
        return dict(F=F[:k], Q=Q[:k], err=err[:k], iters=k)

    def resetHistory(self) -> None:
        check(self._L.psx_scft_reset_history(self._h))

    def boxStress(self) -> np.ndarray:
        out = np.zeros(3)
        check(self._L.psx_scft_box_stress(self._h, _ptr(out)))
        return out

    def backward(self, g_phiA=None, g_phiB=None, g_Q: float = 0.0, box: bool = True):
        """Reverse mode of (w_A, w_B, box) -> (phi_A, phi_B, Q) at the current
        fields (after solveQ): for L = <g_phiA, phi_A> + <g_phiB, phi_B> +
        g_Q Q, returns (dL/dw_A, dL/dw_B, dL/dL_xyz or None)."""
        a = None if g_phiA is None else _f64(g_phiA).reshape(-1)
        b = None if g_phiB is None else _f64(g_phiB).reshape(-1)
        for x in (a, b):
            if x is not None and x.size != self.N:
                raise PsxError(2, "cotangent size differs from the grid")
        gA, gB = np.empty(self.N), np.empty(self.N)
        gL = np.zeros(3) if box else None
        check(self._L.psx_scft_backward_host(self._h, None if a is None else _ptr(a),
                                             None if b is None else _ptr(b), float(g_Q),
                                             _ptr(gA), _ptr(gB), None if gL is None else _ptr(gL)))
        return gA.reshape(self.shape), gB.reshape(self.shape), gL

    def setQuadrature(self, mode) -> None:
        """'reference' (Simpson 0.33333, default) or 'trapezoid' (split-step
        consistent: the discrete F is stationary at the fixed point)."""
        m = {"reference": 0, "simpson": 0, "trapezoid": 1, "consistent": 1}.get(mode, mode)
        check(self._L.psx_scft_set_quadrature(self._h, int(m)))

    def setQuadKernel(self, kind: str) -> None:
        """'auto' (default: small FP32 grids split each cell's contour sum
        over a block's waves) or 'serial' (every cell summed in slice order)."""
        k = {"auto": 0, "serial": 1}.get(kind, kind)
        check(self._L.psx_scft_set_quad_kernel(self._h, int(k)))

    def setCheckpointing(self, every: int) -> None:
        """Keep only every `every`-th backward slice (0: store all slices);
        densities, iterations and the stress recompute the rest chunk by
        chunk.  Linear chains only."""
        check(self._L.psx_scft_set_checkpointing(self._h, int(every)))

    def setCell(self, cell) -> None:
        """New cell sizes (box / grid); fields are kept."""
        self.cell = tuple(float(c) for c in cell)
        check(self._L.psx_scft_set_cell(self._h, *self.cell))

    def paramGrad(self):
        """(dF/dchiN, dF/dfA) at fixed fields (after update() or iterate())."""
        a, b = ctypes.c_double(), ctypes.c_double()
        check(self._L.psx_scft_param_grad(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def timeStep(self, reps: int = 20):
        """(ms per dual contour step, [ms per z, y, x pass])."""
        step = ctypes.c_double()
        passes = (ctypes.c_double * 3)()
        check(self._L.psx_scft_time_step(self._h, int(reps), ctypes.byref(step),
                                         ctypes.cast(passes, _dp)))
        return step.value, list(passes)

    def setVariant(self, variant: int, passes=None) -> None:
        """Line kernel of the given passes (all by default): -1 default,
        1 LDS Stockham, 2 / 3 alternative register splits, 4 128-byte LDS
        row tiles (strided passes), 5 lane pairs 8x16 (n = 128)."""
        na = sum(1 for n in self.shape if n > 1) or 1
        for p in (range(na) if passes is None else passes):
            check(self._L.psx_scft_set_variant(self._h, int(p), int(variant)))

    def setSweepMode(self, mode: str) -> None:
        """"lines": two line launches per contour step; "planes": one launch
        per step, the z and y round trips of each x-plane staged in LDS
        (3D grids with ny == nz in {32, 48, 64, 96, 128}; the default there).
        Fields are kept; propagate again after switching."""
        check(self._L.psx_scft_set_sweep_mode(self._h, {"lines": 0, "planes": 1}[mode]))

    def sweepMode(self) -> tuple:
        """(current mode, whether plane launches are available)."""
        m, a = ctypes.c_int(), ctypes.c_int()
        check(self._L.psx_scft_sweep_mode(self._h, ctypes.byref(m), ctypes.byref(a)))
        return ("planes" if m.value else "lines"), bool(a.value)

    def autotune(self, reps: int = 10) -> list:
        """Pick the fastest line kernel per pass on this engine's data, then
        the faster sweep mode where planes are available; returns the
        choice per pass (the variants of setVariant)."""
        ch = (ctypes.c_int * 3)()
        check(self._L.psx_scft_autotune(self._h, int(reps), ch))
        return list(ch)

    @classmethod
    def _borrow(cls, handle, shape, cell, fA, ds, chiN):
        """View of an engine owned by someone else (a batch)."""
        self = cls.__new__(cls)
        self._L = _abi.load()
        self._h = handle
        self._owned = False
        self.shape = tuple(int(x) for x in shape)
        self.cell = tuple(float(c) for c in cell)
        self.N = int(np.prod(self.shape))
        self.fA, self.ds, self.chiN = float(fA), float(ds), float(chiN)
        self._steps()
        return self

    def close(self):
        if getattr(self, "_h", None) and getattr(self, "_owned", True):
            self._L.psx_scft_destroy(self._h)
        self._h = None

    __del__ = close


@dataclass
class Block:
    """One ``Block`` of a ``blockCopolymer`` (pspolymer/PsBlockCopolymer.cpp,
    psbase/PsBlockBase.cpp:110-131): species "A"/"B" (or 0/1; the block's
    scfield), lengthfrac, bSegRatio, and the indices of the blocks named in
    headjoined / tailjoined (empty = freeEnd)."""
    species: object
    lengthfrac: float
    bSegRatio: float = 1.0
    head: tuple = ()
    tail: tuple = ()

    def c(self) -> _abi.BlockDesc:
        sp = {"A": 0, "B": 1, 0: 0, 1: 1}.get(self.species)
        if sp is None:
            raise ValueError(f"species must be A or B, not {self.species!r}")
        if len(self.head) > _abi.PSX_MAX_JOIN or len(self.tail) > _abi.PSX_MAX_JOIN:
            raise ValueError("too many joined blocks")
        d = _abi.BlockDesc()
        d.species, d.lengthfrac, d.bSegRatio = sp, float(self.lengthfrac), float(self.bSegRatio)
        d.nhead, d.ntail = len(self.head), len(self.tail)
        for i, j in enumerate(self.head):
            d.head[i] = int(j)
        for i, j in enumerate(self.tail):
            d.tail[i] = int(j)
        return d


class BlockCopolymerSCFT(DiblockSCFT):
    """Bulk melt of one AB block copolymer of any tree architecture
    (``PsBlockCopolymer`` with ``flexPseudoSpec`` blocks sharing one ds):
    linear multiblocks, stars, combs.  Same methods as ``DiblockSCFT``;
    ``fA`` is the total A fraction."""

    def __init__(self, shape, cell, blocks, ds, chiN, prec="fp64", stream=None,
                 quadrature="reference"):
        self._L = _abi.load()
        self.shape = tuple(int(s) for s in shape)
        self.cell = tuple(float(c) for c in cell)
        self.N = int(np.prod(self.shape))
        self.blocks = [b if isinstance(b, Block) else Block(*b) for b in blocks]
        self.ds, self.chiN = float(ds), float(chiN)
        self.fA = sum(b.lengthfrac for b in self.blocks if b.species in ("A", 0))
        g = _abi.Grid(*self.shape, *self.cell)
        arr = (_abi.BlockDesc * len(self.blocks))(*[b.c() for b in self.blocks])
        h = ctypes.c_void_p()
        check(self._L.psx_scft_create_blocks(ctypes.byref(g), len(self.blocks), arr, self.ds,
                                             self.chiN, _prec(prec), stream, ctypes.byref(h)))
        self._h = h
        self._steps()
        if quadrature != "reference":
            self.setQuadrature(quadrature)


@dataclass
class Polymer:
    """One ``<Polymer>`` of a blend (kind blockCopolymer,
    pspolymer/PsPolymerHldr.cpp:35-83): volfrac, length (the first polymer's
    length is the scale length) and its ``Block``s, joined by indices
    within the polymer."""
    volfrac: float
    length: float
    blocks: tuple


class BlendSCFT(DiblockSCFT):
    """Bulk melt of several AB block copolymers sharing the A and B fields
    (e.g. a diblock with a homopolymer, the reference's
    pstests/tests/diblock/abHomopMix.pre): each polymer has its own Q and
    contributes volfrac / lengthRatio of its qqt to the densities.  Same
    methods as ``DiblockSCFT``; ``calcBigQ`` is the first polymer's Q and
    ``polymerQ`` gives all of them."""

    def __init__(self, shape, cell, polymers, ds, chiN, prec="fp64", stream=None,
                 quadrature="reference"):
        self._L = _abi.load()
        self.shape = tuple(int(s) for s in shape)
        self.cell = tuple(float(c) for c in cell)
        self.N = int(np.prod(self.shape))
        self.polymers = [p if isinstance(p, Polymer) else Polymer(*p) for p in polymers]
        for p in self.polymers:
            p.blocks = tuple(b if isinstance(b, Block) else Block(*b) for b in p.blocks)
        self.blocks = [b for p in self.polymers for b in p.blocks]
        self.ds, self.chiN = float(ds), float(chiN)
        self.fA = sum(p.volfrac * b.lengthfrac for p in self.polymers for b in p.blocks
                      if b.species in ("A", 0))
        g = _abi.Grid(*self.shape, *self.cell)
        parr = (_abi.PolymerDesc * len(self.polymers))(
            *[_abi.PolymerDesc(float(p.volfrac), float(p.length), len(p.blocks)) for p in self.polymers])
        arr = (_abi.BlockDesc * len(self.blocks))(*[b.c() for b in self.blocks])
        h = ctypes.c_void_p()
        check(self._L.psx_scft_create_blend(ctypes.byref(g), len(self.polymers), parr, arr, self.ds,
                                            self.chiN, _prec(prec), stream, ctypes.byref(h)))
        self._h = h
        self._steps()
        if quadrature != "reference":
            self.setQuadrature(quadrature)

    def polymerQ(self) -> np.ndarray:
        """Q of every polymer, in polymer order."""
        out = np.zeros(_abi.PSX_MAX_POLYMERS)
        n = ctypes.c_int()
        check(self._L.psx_scft_polymer_Q(self._h, _ptr(out), ctypes.byref(n)))
        return out[:n.value].copy()


class BatchSCFT:
    """Independent diblock instances (parameter sweep) advanced together.

    specs: list of dicts with shape, cell, fA, ds, chiN (and optional bA, bB).
    All instances need the same contour step count (same ds)."""

    def __init__(self, specs, prec="fp32", stream=None):
        self._L = _abi.load()
        n = len(specs)
        grids = (_abi.Grid * n)()
        chains = (_abi.Diblock * n)()
        for i, sp in enumerate(specs):
            grids[i] = _abi.Grid(*[int(x) for x in sp["shape"]], *[float(c) for c in sp["cell"]])
            chains[i] = _abi.Diblock(float(sp["fA"]), float(sp["ds"]), float(sp.get("bA", 1.0)),
                                     float(sp.get("bB", 1.0)), float(sp["chiN"]))
        h = ctypes.c_void_p()
        check(self._L.psx_batch_create(n, grids, chains, _prec(prec), stream, ctypes.byref(h)))
        self._h = h
        self.specs = [dict(sp) for sp in specs]
        nc = ctypes.c_int()
        check(self._L.psx_batch_size(h, None, ctypes.byref(nc)))
        self.nclasses = nc.value
        self.instances = []
        for i, sp in enumerate(specs):
            eh = ctypes.c_void_p()
            check(self._L.psx_batch_instance(h, i, ctypes.byref(eh)))
            self.instances.append(DiblockSCFT._borrow(eh, sp["shape"], sp["cell"], sp["fA"],
                                                      sp["ds"], sp["chiN"]))

    def __len__(self):
        return len(self.instances)

    def setPrecision(self, prec) -> None:
        check(self._L.psx_batch_set_precision(self._h, _prec(prec)))

    def setSweepMode(self, cls: int, mode: str) -> None:
        """Sweep mode ("lines" / "planes") of size class cls."""
        check(self._L.psx_batch_set_sweep_mode(self._h, int(cls), {"lines": 0, "planes": 1}[mode]))

    def sweepMode(self, cls: int) -> str:
        m = ctypes.c_int()
        check(self._L.psx_batch_sweep_mode(self._h, int(cls), ctypes.byref(m)))
        return "planes" if m.value else "lines"

    def autotune(self, reps: int = 10):
        """Per grid size and pass: fastest line kernel (-1 register default,
        1 Stockham, 2 alternative register split), then per class the
        faster sweep mode where planes are available (see sweepMode)."""
        ch = (ctypes.c_int * (3 * self.nclasses))()
        check(self._L.psx_batch_autotune(self._h, int(reps), ch))
        return [list(ch[3 * c:3 * c + 3]) for c in range(self.nclasses)]

    def iterate(self, n_iters: int, mixing: Mixing | None = None, use_graph: bool = True) -> dict:
        mixing = mixing or Mixing()
        n = len(self.instances)
        F = np.zeros(n * n_iters)
        err = np.zeros(n * n_iters)
        done = ctypes.c_int()
        m = mixing.c()
        check(self._L.psx_batch_iterate(self._h, ctypes.byref(m), int(n_iters), int(bool(use_graph)),
                                        _ptr(F), _ptr(err), ctypes.byref(done)))
        return dict(F=F.reshape(n, n_iters), err=err.reshape(n, n_iters), iters=done.value)

    def close(self):
        for inst in getattr(self, "instances", []):
            inst._h = None
        if getattr(self, "_h", None):
            self._L.psx_batch_destroy(self._h)
            self._h = None

    __del__ = close
