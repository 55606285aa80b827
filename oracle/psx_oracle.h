/*
 * psx_oracle.h -- CPU oracle for the SCFT hot path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine
 * is compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * It restates, in plain single-file C++, the reference algorithm of
 * PolySwift++ for one bulk block-copolymer SCFT iteration:
 *   - k^2 table and exp(-ds k^2 b^2)      psbase/PsGridField.cpp:66-104,
 *                                          pspolymer/PsFlexPseudoSpec.cpp:188-206
 *   - scaledFFTPair (unnormalised pair)   psfft/PsFFTW.cpp:196-224
 *   - solveQ split-step propagator        pspolymer/PsFlexPseudoSpec.cpp:244-293
 *   - w-factor                            pspolymer/PsFlexPseudoSpec.cpp:159-181
 *   - block graph for a linear diblock    pspolymer/PsBlockCopolymer.cpp:189-313
 *   - Q                                   pspolymer/PsBlock.cpp:205-237
 *   - Simpson(0.33333)+trapezoid quadrature pspolymer/PsBlock.cpp:256-380
 *   - densities                           pspolymer/PsBlock.cpp:154-196
 *   - block steps                         psbase/PsBlockBase.cpp:252-326
 *   - steepest-descent field update       pseffhamil/PsSteepDUpdater.cpp:159-333
 *   - Flory dF/dphi and F terms           pseffhamil/PsFloryInteraction.cpp:204-321
 *   - pressure shift                      psphysf/PsConstraint.cpp:184-262
 *   - free energy                         pseffhamil/PsPolymerUpdater.cpp:173-341
 *   - iteration order                     pstrol/PsDomain.cpp:332-372
 * plus two components the reference does not have (Anderson mixing and the
 * analytic box stress), restated from their published definitions.
 *
 * The FFT is this file's own recursive mixed-radix Cooley-Tukey transform
 * (FFTW 2.x is not available); it shares no code with the GPU Stockham FFT.
 */
#ifndef PSX_ORACLE_H
#define PSX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One block of a general block graph (pspolymer/PsBlockCopolymer.cpp,
 * psbase/PsBlockBase.cpp:110-131,205-240): species 0 = A, 1 = B (the Block's
 * scfield), lengthfrac, bSegRatio, and the indices of the blocks joined at
 * its head and tail (headjoined / tailjoined; none = freeEnd). */
#define ORC_MAX_JOIN 8
typedef struct orc_block {
  int species;
  double frac;
  double b;
  int nhead, ntail;
  int head[ORC_MAX_JOIN], tail[ORC_MAX_JOIN];
} orc_block;

/* One polymer of a blend (a <Polymer> of the deck, pspolymer/PsPolymerHldr.cpp:
 * 69-83): volfrac, lengthRatio = length / the first polymer's length, and
 * its blocks: the next nblocks entries of orc_params.blocks, whose
 * head/tail indices count within the polymer. */
#define ORC_MAX_POLY 8
typedef struct orc_polymer {
  double volfrac;
  double lratio;
  int nblocks;
} orc_polymer;

/* Run parameters of one AB melt (bulk, no walls, no solvent): the diblock of
 * fA, bA, bB, or the block graph `blocks` when nblocks > 0, or the blend of
 * the npoly polymers `polys` (blocks then holds all their blocks in order). */
typedef struct orc_params {
  int nx, ny, nz;        /* grid cells; nz = 1 (and ny = 1) for 2D (1D) */
  double dx, dy, dz;     /* cell sizes in Rg (reference "cellSizes") */
  double ds;             /* contour step (both blocks) */
  double fA;             /* lengthfrac of block A; block B has 1 - fA */
  double chiN;           /* Flory chi times N */
  double bA, bB;         /* segment-length ratios (reference bSegRatio) */
  int mix;               /* 0 steepest descent (lambda0, lambda1); 1 Anderson */
  double lambda0, lambda1;
  int depth;             /* Anderson history depth m */
  double beta;           /* Anderson mixing parameter */
  double anderson_start; /* use Picard while residual RMS > this */
  int f_order;           /* 0: F after the w update (reference); 1: consistent */
  int quad;              /* 0: Simpson 0.33333 (reference); 1: trapezoid */
  int nblocks;           /* 0: diblock of fA; > 0: the block graph below */
  const orc_block* blocks;
  int npoly;             /* 0: one polymer of volfrac = lengthRatio = 1 */
  const orc_polymer* polys;
} orc_params;

/* Thread count used by the FFT line loops and element-wise loops (OpenMP). */
void orc_set_threads(int n);
int orc_get_threads(void);

/* Block steps per psbase/PsBlockBase.cpp:252-326; returns -1 on error. */
int orc_block_steps(double length_frac, double ds);

/* kfac[i,j,k] = exp(-ds*b^2*k2), full grid, C order (k fastest). */
void orc_kfac(int nx, int ny, int nz, double dx, double dy, double dz,
              double ds, double b, double* kfac);

/* In-place unnormalised 3D complex DFT; sign -1 forward, +1 backward. */
int orc_fft3(int nx, int ny, int nz, double* re, double* im, int sign);
/* O(N^2) direct 3D DFT (exact-index twiddles), for tiny grids only. */
int orc_dft3_brute(int nx, int ny, int nz, const double* re, const double* im,
                   double* ore, double* oim, int sign);

/* res = Re(IFFT_u(kdata .* FFT_u(data + 0i))), no 1/N (psfft/PsFFTW.cpp:196-224). */
int orc_scaled_fft_pair(int nx, int ny, int nz, const double* data,
                        const double* kdata, double* res);

/* One block sweep: slices[0] = q0, slices[s] = split step of slices[s-1],
 * s = 1..steps, with wfac = exp(-ds w/2) (pspolymer/PsFlexPseudoSpec.cpp:244-293). */
int orc_solve_q(int nx, int ny, int nz, const double* kfac, const double* wfac,
                const double* q0, int steps, double* slices);

/* qqt = (1/Q) * sum_j c_j qs[j] qts[M-1-j], Simpson 0.33333 + trapezoid tail
 * (pspolymer/PsBlock.cpp:256-380).  qs, qts hold M = steps+1 slices. */
void orc_quadrature(long n, int steps, double ds, const double* qs,
                    const double* qts, double Q, double* out);

/* Trapezoid contour rule ds [1/2, 1, ..., 1, 1/2] (no reference counterpart:
 * the exact derivative of the split-step Q with respect to w). */
void orc_quadrature_trap(long n, int steps, double ds, const double* qs,
                         const double* qts, double Q, double* out);

/* Full propagation of a diblock: fills densities and Q.  If slices != NULL it
 * receives qA[0..sA], qB[0..sB], qtB[0..sB], qtA[0..sA] in that order. */
int orc_diblock_densities(const orc_params* P, const double* wA,
                          const double* wB, double* phiA, double* phiB,
                          double* Q, double* slices);

/* Full propagation of a block graph (pspolymer/PsBlockCopolymer.cpp:189-313,
 * pspolymer/PsBlock.cpp:146-147,154-237,387-472): fills the species densities
 * and Q (for a blend, one Q per polymer; each block's density is scaled by
 * its polymer's volfrac / lengthRatio, pspolymer/PsBlock.cpp:170-173).  If slices != NULL it receives, block by block, the M = steps+1
 * slices propagated from the head (index from the head) followed by the M
 * slices propagated from the tail (index from the tail).  Returns -1 on a bad
 * graph (block steps, indices, a cycle), -2 on NaN Q.  Works for the diblock
 * (nblocks = 0) too, with blocks A = 0 and B = 1. */
int orc_graph_densities(const orc_params* P, const double* wA,
                        const double* wB, double* phiA, double* phiB,
                        double* Q, double* slices);

/* Total slices that orc_graph_densities writes (in fields of N), or -1. */
long orc_graph_slice_count(const orc_params* P);

/* Species totals f_A, f_B (sum of lengthfrac * volfrac over the species'
 * blocks, psbase/PsBlockBase.cpp:185-189). */
void orc_species_fracs(const orc_params* P, double* fA, double* fB);

/* One steepest-descent update + pressure shift + F, in reference order
 * (pseffhamil/PsSteepDUpdater.cpp:159-203).  Updates wA, wB, p in place. */
double orc_sd_update(const orc_params* P, long n, double* wA, double* wB,
                     double* p, const double* phiA, const double* phiB,
                     double Q);

/* Free energy F for given fields (pseffhamil/PsPolymerUpdater.cpp:173-226). */
double orc_free_energy(const orc_params* P, long n, const double* wA,
                       const double* wB, const double* phiA,
                       const double* phiB, double Q);
/* The same for a blend: Qs holds one Q per polymer, and the chain term is
 * -sum_p volfrac_p ln Q_p / lengthRatio_p (PsPolymerUpdater.cpp:198-207). */
double orc_free_energy_poly(const orc_params* P, long n, const double* wA,
                            const double* wB, const double* phiA,
                            const double* phiB, const double* Qs);
/* Number of polymers (1 unless npoly > 1). */
int orc_polymer_count(const orc_params* P);

/* Run SCFT iterations until n_iters or residual RMS < tol (tol <= 0: never).
 * Histories (length n_iters, may be NULL): F, Q, residual RMS.
 * Returns the number of iterations done, or -1 on error. */
int orc_scft_run(const orc_params* P, double* wA, double* wB, double* p,
                 int n_iters, double tol, double* F_hist, double* Q_hist,
                 double* err_hist, double* phiA, double* phiB);

/* The same iterations timed: t_hist[it] = wall-clock seconds of iteration it
 * (std::chrono::steady_clock around propagation, Q, densities, the field
 * update and F, as the reference's TxuTimer wraps PsDomain::update,
 * polyswift/polyswift.cxx:571-574).  Returns the iterations done. */
int orc_scft_run_timed(const orc_params* P, double* wA, double* wB, double* p,
                       int n_iters, double* F_hist, double* t_hist);

/* Analytic box stress dF/dL_d (d = x,y,z) at fixed w for the diblock: the
 * derivative of -ln Q through k^2 only (fields held fixed in cell units). */
int orc_box_stress(const orc_params* P, const double* wA, const double* wB,
                   double* dFdL);

/* Time `steps` faithful contour steps on an nx*ny*nz grid (seconds). */
double orc_time_steps(int nx, int ny, int nz, int steps);

/* k^2 of PsGridField::calck2 (psbase/PsGridField.cpp:66-104), C order. */
void orc_k2(int nx, int ny, int nz, double dx, double dy, double dz, double* k2);
/* The rest of PsFFTBase (psfft/PsFFTW.cpp:121-296), unnormalised:
 * |FFT(data)|^2; Re(IFFT(FFT(d1) FFT(d2))); Re(IFFT(kdata FFT(i data)));
 * Re(FFT(data)); Re(IFFT(data)). */
int orc_fft_forward_abs(int nx, int ny, int nz, const double* data, double* res);
int orc_fft_convolve_re(int nx, int ny, int nz, const double* d1, const double* d2, double* res);
int orc_fft_scaled_pair_im(int nx, int ny, int nz, const double* data, const double* kdata,
                           double* res);
int orc_fft_forward_re(int nx, int ny, int nz, const double* data, double* res);
int orc_fft_backward_re(int nx, int ny, int nz, const double* data, double* res);
/* Spectral filter in place (PsSimpleSpecFilter for cells == NULL, else
 * PsMultiSpecFilter with spectral cells cells[3]); -2 if a cell size does
 * not divide its grid size. */
int orc_spec_filter(int nx, int ny, int nz, double* w, double cutoff, double strength,
                    const int* cells, double vwall);
/* The reference's random stream (PsRandom: MT19937 with init_genrand
 * seeding, real() = y / 2^32, Marsaglia polar Gaussians). */
typedef struct orc_rng orc_rng;
orc_rng* orc_rng_create(unsigned long seed);
void orc_rng_destroy(orc_rng* g);
void orc_rng_raw32(orc_rng* g, long n, unsigned int* out);
void orc_rng_uniform(orc_rng* g, long n, double* out);
void orc_rng_gaussian(orc_rng* g, long n, double* out);
/* Steepest descent (P->mix == 0) with the noise term drawn from g. */
int orc_scft_run_noise(const orc_params* P, double* wA, double* wB, double* p, int n_iters,
                       double* F_hist, double* Q_hist, double* err_hist, double strength,
                       orc_rng* g);
/* PsPoissonUpdater::update in place on charge and psi. */
int orc_poisson_update(int nx, int ny, int nz, double dx, double dy, double dz, double lB,
                       double* charge, double* psi);

#ifdef __cplusplus
}
#endif
#endif
