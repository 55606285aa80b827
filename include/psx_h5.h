/*
 * psx_h5.h -- C ABI of libpsx_h5.so: the PhysField dump/restore file format
 * of PolySwift++ (HDF5, VizSchema attributes), host side.
 *
 * One file per PhysField per dump, named <prefix>_<field>_<1000000+step>.h5
 * by the caller (TxIoBase::getDumpFileName as used in
 * psphysf/PsPhysFldFuncs.cpp:125,153; psutils/python/ps-run.py.in:37-50).
 * It holds two FP64 datasets in C order: the field's density
 * ("MonomerDensity") and its conjugate ("ConjugateDensity" for a monomer
 * density, "PressureField" for the constraint; psphysf/PsMonomerDens.cpp:
 * 23-26, psphysf/PsConstraint.cpp:22-26).  Their shape is (nx, ny, 1) when
 * nz == 1 and (nx, ny, nz, 1) otherwise (PsPhysFldFuncs::buildSolvers,
 * psphysf/PsPhysFldFuncs.cpp:88-113).
 *
 * Plain C; every call returns 0 on success and a negative code on failure,
 * with psx_h5_last_error() giving a thread-local message.
 */
#ifndef PSX_H5_H
#define PSX_H5_H

#ifdef __cplusplus
extern "C" {
#endif

/* Write one PhysField dump (PsPhysFldFuncs::dump, dumpField,
 * appendMetaDataset, appendMetaGroup; psphysf/PsPhysFldFuncs.cpp:119-143,
 * 214-350): the two datasets, their VizSchema attributes (vsType variable,
 * vsMesh <field>Global, vsLimits <field>GlobalLimits, GridSizes, vsCentering
 * zonal, vsLabels, vsAxisLabels, vsTimeGroup time) and the groups
 * <field>Global (uniform mesh), <field>GlobalLimits (Cartesian limits) and
 * time (vsStep, vsTime).  An existing file is not overwritten (the
 * reference skips it too): returns 1 without writing. */
int psx_h5_write_physfield(const char* path, const char* field, int nx, int ny, int nz,
                           const double* cell, const char* dens_name, const double* dens,
                           const char* conj_name, const double* conj, long long step);

/* Read the two datasets of a PhysField dump (PsPhysFldFuncs::restore,
 * readField; psphysf/PsPhysFldFuncs.cpp:148-164,234-248) into nx*ny*nz
 * FP64 arrays (either output may be NULL).  Fails if a dataset is missing or
 * its element count differs. */
int psx_h5_read_physfield(const char* path, int nx, int ny, int nz,
                          const char* dens_name, double* dens,
                          const char* conj_name, double* conj);

/* Last error message of the calling thread. */
void psx_h5_last_error(char* buf, unsigned long len);

#ifdef __cplusplus
}
#endif
#endif
