"""PhysField checkpoint files: the reference's HDF5 dump / restore format.

One file per PhysField per dump, ``<prefix>_<field>_<1000000 + step>.h5``
(the numbering psutils/python/ps-run.py.in:37-50 reads back), written by
``libpsx_h5.so`` (include/psx_h5.h):

* a monomer-density field (``PsMonomerDens``) holds ``MonomerDensity`` (phi)
  and ``ConjugateDensity`` (w), psphysf/PsMonomerDens.cpp:23-26;
* the constraint field (``PsConstraint``, e.g. ``defaultPressure``) holds
  ``MonomerDensity`` (the wall density, 0 in bulk) and ``PressureField`` (p),
  psphysf/PsConstraint.cpp:22-26.

``dump`` mirrors ``PsDomain::dump`` for a melt's three PhysFields and
``restore`` mirrors ``PsDomain::restore`` (pstrol/PsDomain.cpp:306-396):
it sets w_A, w_B and p of an engine from the files of one dump step.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
H5LIB_PATH = os.path.join(HERE, "libpsx_h5.so")
_lib = None
_dp = ctypes.POINTER(ctypes.c_double)

DEFAULT_FIELDS = ("totA", "totB", "defaultPressure")


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(H5LIB_PATH):
            raise RuntimeError(f"{H5LIB_PATH} is missing: build it with "
                               "`python -m polyswift_plusplus_amd.build_ext` (needs an HDF5 library)")
        L = ctypes.CDLL(H5LIB_PATH)
        L.psx_h5_write_physfield.restype = ctypes.c_int
        L.psx_h5_write_physfield.argtypes = [ctypes.c_char_p, ctypes.c_char_p] + [ctypes.c_int] * 3 + [
            _dp, ctypes.c_char_p, _dp, ctypes.c_char_p, _dp, ctypes.c_longlong]
        L.psx_h5_read_physfield.restype = ctypes.c_int
        L.psx_h5_read_physfield.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 3 + [
            ctypes.c_char_p, _dp, ctypes.c_char_p, _dp]
        L.psx_h5_last_error.restype = None
        L.psx_h5_last_error.argtypes = [ctypes.c_char_p, ctypes.c_ulong]
        _lib = L
    return _lib


def _err() -> str:
    buf = ctypes.create_string_buffer(1024)
    _load().psx_h5_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def _arr(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if a.size != n:
        raise ValueError(f"field has {a.size} values, grid has {n}")
    return a


def dump_name(prefix: str, field: str, step: int) -> str:
    return f"{prefix}_{field}_{1000000 + int(step)}.h5"


def write_physfield(path, field, shape, cell, dens_name, dens, conj_name, conj, step) -> bool:
    """Write one PhysField file; False if it already existed (not overwritten)."""
    n = int(np.prod(shape))
    d, c = _arr(dens, n), _arr(conj, n)
    h = (ctypes.c_double * 3)(*[float(x) for x in cell])
    rc = _load().psx_h5_write_physfield(path.encode(), field.encode(), *[int(x) for x in shape],
                                        h, dens_name.encode(), d.ctypes.data_as(_dp),
                                        conj_name.encode(), c.ctypes.data_as(_dp), int(step))
    if rc < 0:
        raise OSError(f"{path}: {_err()}")
    return rc == 0


def read_physfield(path, shape, dens_name, conj_name):
    """(density, conjugate) arrays of one PhysField file, shaped like the grid."""
    n = int(np.prod(shape))
    d, c = np.empty(n), np.empty(n)
    rc = _load().psx_h5_read_physfield(path.encode(), *[int(x) for x in shape], dens_name.encode(),
                                       d.ctypes.data_as(_dp), conj_name.encode(), c.ctypes.data_as(_dp))
    if rc < 0:
        raise OSError(f"{path}: {_err()}")
    return d.reshape(shape), c.reshape(shape)


def dump(engine, prefix: str, step: int, fields=DEFAULT_FIELDS) -> list:
    """Write the melt's PhysFields at `step`: A and B monomer densities with
    their conjugate fields, and the pressure constraint.  Densities are those
    of the last densities()/iterate() call (zeros if none yet)."""
    wA, wB, p = engine.getFields()
    try:
        phiA, phiB = engine.getDensities()
    except Exception:
        phiA, phiB = np.zeros(engine.shape), np.zeros(engine.shape)
    a, b, c = fields
    out = []
    for name, dens, dname, conj, cname in ((a, phiA, "MonomerDensity", wA, "ConjugateDensity"),
                                           (b, phiB, "MonomerDensity", wB, "ConjugateDensity"),
                                           (c, np.zeros(engine.shape), "MonomerDensity", p, "PressureField")):
        path = dump_name(prefix, name, step)
        write_physfield(path, name, engine.shape, engine.cell, dname, dens, cname, conj, step)
        out.append(path)
    return out


def restore(engine, prefix: str, step: int, fields=DEFAULT_FIELDS) -> None:
    """Set w_A, w_B and p of the engine from the files of dump `step`."""
    a, b, c = fields
    _, wA = read_physfield(dump_name(prefix, a, step), engine.shape, "MonomerDensity", "ConjugateDensity")
    _, wB = read_physfield(dump_name(prefix, b, step), engine.shape, "MonomerDensity", "ConjugateDensity")
    _, p = read_physfield(dump_name(prefix, c, step), engine.shape, "MonomerDensity", "PressureField")
    engine.setFields(wA, wB, p)
