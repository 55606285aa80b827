/*
 * psx_scft.h -- C ABI of the MI355X-native SCFT engine (libpsx_scft.so).
 *
 * Plain C: pointers, sizes and status codes.  No C++ exception crosses this
 * boundary; every call returns psx_status (0 = OK) and psx_last_error()
 * returns a thread-local message for the last failing call on that thread.
 * Device pointers ("d_") live in HBM of the current HIP device; host
 * pointers ("h_") are ordinary host memory.  Streams are hipStream_t passed
 * as void* (NULL = the object creates and owns a default-priority stream).
 *
 * Each entry point cites the reference interface it replaces
 * (paths relative to the PolySwift++ tree polyswift/).
 */
#ifndef PSX_SCFT_H
#define PSX_SCFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int psx_status;
enum {
  PSX_OK = 0,
  PSX_ERR_INVALID = 1,   /* bad argument (null pointer, bad enum, ...) */
  PSX_ERR_SHAPE = 2,     /* grid shape / size mismatch */
  PSX_ERR_STEPS = 3,     /* non-integral or too few block steps */
  PSX_ERR_NAN = 4,       /* Q is NaN/Inf (reference PsBlock::calcBigQ) */
  PSX_ERR_HIP = 5,       /* HIP runtime error */
  PSX_ERR_NOMEM = 6,     /* device allocation failed */
  PSX_ERR_STATE = 7      /* call out of order (e.g. Q before propagate) */
};

/* Arithmetic type of the contour sweep (FFTs, w-factors, stored slices).
 * Q, the contour quadrature, the field update and F are always FP64. */
typedef enum { PSX_FP64 = 0, PSX_FP32 = 1 } psx_prec;

/* Copies the last error message of the calling thread into buf. */
void psx_last_error(char* buf, size_t len);
/* Library version string. */
const char* psx_version(void);

/* ===================================================================== */
/* FFT pair plugin.  Replaces PsFFTBase<T,NDIM> (psbase/PsFFTBase.h:39-147) */
/* as implemented by PsFFTW (psfft/PsFFTW.cpp:53-224).                   */
/* ===================================================================== */
typedef struct psx_fft psx_fft;

/* Plan for an nx*ny*nz C-order real grid (PsFFTW::buildData,
 * psfft/PsFFTW.cpp:53-88).  nz = 1 (and ny = 1) give 2D (1D) grids. */
psx_status psx_fft_plan_create(int nx, int ny, int nz, psx_prec prec,
                               void* stream, psx_fft** out);
psx_status psx_fft_plan_destroy(psx_fft* plan);
/* Number of real elements of data/kdata/res (PsFFTBase::getFFTSize,
 * psbase/PsFFTBase.h:75; for the serial plan this is nx*ny*nz). */
size_t psx_fft_size(const psx_fft* plan);
/* res = Re(IFFT(kdata .* FFT(data + 0i))), both transforms unnormalised,
 * forward sign -, backward +; kdata full grid, C order, real
 * (PsFFTBase::scaledFFTPair, psbase/PsFFTBase.h:105-106;
 *  psfft/PsFFTW.cpp:196-224).  Device pointers; element type per prec.
 * data/res may alias.  Asynchronous on the plan's stream. */
psx_status psx_fft_scaled_pair(psx_fft* plan, const void* d_data,
                               const void* d_kdata, void* d_res);
/* Same with host arrays (copies over PCIe; synchronous). */
psx_status psx_fft_scaled_pair_host(psx_fft* plan, const void* h_data,
                                    const void* h_kdata, void* h_res);

/* The same pair without a plan and without allocating anything.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc).  The twiddle tables and the complex values between the
 * passes live in one static device array per GPU of PSX_NOALLOC_SCRATCH_BYTES,
 * held until the process ends; grids need every dimension <= PSX_NOALLOC_MAX_LINE
 * and nx*ny*nz*2*sizeof(T) <= PSX_NOALLOC_WORK_BYTES (128^3 in FP64, 2^22
 * cells in FP32), else PSX_ERR_SHAPE.  stream NULL is the null stream.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialised by the caller.
 * Results equal psx_fft_scaled_pair's to rounding (the twiddles are computed
 * on the device). */
#define PSX_NOALLOC_MAX_LINE 1024
#define PSX_NOALLOC_WORK_BYTES (32u * 1024u * 1024u)
#define PSX_NOALLOC_SCRATCH_BYTES (PSX_NOALLOC_WORK_BYTES + 3u * 2u * PSX_NOALLOC_MAX_LINE * 8u)
psx_status psx_fft_scaled_pair_noalloc(int nx, int ny, int nz, psx_prec prec,
                                       void* stream, const void* d_data,
                                       const void* d_kdata, void* d_res);

/* The rest of PsFFTBase (psbase/PsFFTBase.h:83-139, psfft/PsFFTW.cpp:
 * 121-194,226-296): device pointers of N elements of the plan's type,
 * asynchronous on the plan's stream, unnormalised transforms.
 *   forward_abs:    res = |FFT(data)|^2 (the squared modulus, as
 *                   PsFFTW::forwardFFTAbs computes it)
 *   convolve_re:    res = Re(IFFT(FFT(data1) .* FFT(data2)))
 *   scaled_pair_im: res = Re(IFFT(kdata .* FFT(0 + i data)))
 *   forward_re:     res = Re(FFT(data))        (calcForwardFFT)
 *   backward_re:    res = Re(IFFT(data + 0i))  (calcBackwardFFT) */
psx_status psx_fft_forward_abs(psx_fft* plan, const void* d_data, void* d_res);
psx_status psx_fft_convolve_re(psx_fft* plan, const void* d_data1,
                               const void* d_data2, void* d_res);
psx_status psx_fft_scaled_pair_im(psx_fft* plan, const void* d_data,
                                  const void* d_kdata, void* d_res);
psx_status psx_fft_forward_re(psx_fft* plan, const void* d_data, void* d_res);
psx_status psx_fft_backward_re(psx_fft* plan, const void* d_data, void* d_res);

/* Spectral filter of one field, in place (device pointer, plan's type):
 * PsSpecFilterUpdater::subtractAverage (w -= sum(w) / (N - wall_volume)),
 * then the mask filterStrength where |FFT(w)|^2 < cutoffFactor * max, 1
 * elsewhere, applied by the FFT pair, w = pair / N.  spec_cells = NULL:
 * PsSimpleSpecFilter (one maximum, pseffhamil/PsSimpleSpecFilter.cpp:
 * 74-144); spec_cells = {sx, sy, sz}: PsMultiSpecFilter with one maximum
 * per spectral cell (pseffhamil/PsMultiSpecFilter.cpp:123-267, the region
 * map as a serial run lays it out); each s must divide its grid size.
 * cutoff_factor <= 0 leaves w untouched, as the reference returns early. */
psx_status psx_fft_spec_filter(psx_fft* plan, void* d_w, double cutoff_factor,
                               double filter_strength, const int* spec_cells,
                               double wall_volume);
/* PsPoissonUpdater::update for one field (pseffhamil/PsPoissonUpdater.cpp:
 * 214-289), in place: the charge density loses its mean; psi becomes
 * pair(charge, 4 * 3.1415 * bjerrum_len / k^2) / N (k = 0 mode 0, k^2 from
 * the cell sizes as PsGridField::calck2 forms it), except where that changes
 * psi by more than 20%: there psi becomes 1.2 psi_old / N, the reference's
 * limiter as written. */
psx_status psx_fft_poisson_update(psx_fft* plan, void* d_charge, void* d_psi,
                                  double bjerrum_len, double dx, double dy,
                                  double dz);
/* Wait for the plan's stream. */
psx_status psx_fft_synchronize(psx_fft* plan);

/* Device buffers for callers without a HIP runtime of their own (FFI
 * hosts): allocate, free, and synchronous copies.  No reference
 * counterpart (the reference passes host arrays). */
psx_status psx_buffer_create(size_t bytes, void** d_out);
psx_status psx_buffer_destroy(void* d);
psx_status psx_buffer_upload(void* d_dst, const void* h_src, size_t bytes);
psx_status psx_buffer_download(void* h_dst, const void* d_src, size_t bytes);
/* Memory probe (no reference counterpart): allocates a buffer of `bytes`,
 * then times `reps` back-to-back sweeps over it after one warm sweep (mode
 * 0: read-only, 16-byte loads; mode 1: copy into a second buffer of the
 * same size) with hipEvents on the null stream; *h_ms = average ms per
 * sweep.  Repeated sweeps over buffers that fit the L2 or the Infinity
 * Cache show those levels' rates; larger ones, HBM's. */
psx_status psx_probe_stream(size_t bytes, int reps, int mode, double* h_ms);

/* ===================================================================== */
/* SCFT engine for a bulk AB diblock melt: the block propagators          */
/* (PsBlockBase / PsFlexPseudoSpec), Q, densities, Flory interaction and  */
/* the field updaters (PsSteepDUpdater, plus Anderson mixing, which the   */
/* reference does not have).                                              */
/* ===================================================================== */
typedef struct psx_grid {
  int nx, ny, nz;        /* numCellsGlobal */
  double dx, dy, dz;     /* cellSizes in Rg (psgrid/PsUniCartGrid.cpp:30-43) */
} psx_grid;

typedef struct psx_diblock {
  double fA;             /* lengthfrac of block A; block B has 1 - fA */
  double ds;             /* contour step (reference "ds"), s in [0, 1] */
  double bA, bB;         /* bSegRatio of each block */
  double chiN;           /* chi * scaleLength (PsFloryInteraction.cpp:160-161) */
} psx_diblock;

/* One block of a general block copolymer (a Block of a blockCopolymer
 * Polymer, pspolymer/PsBlockCopolymer.cpp:60-150, psbase/PsBlockBase.cpp:
 * 110-131): species 0 = A, 1 = B (the Block's scfield), lengthfrac,
 * bSegRatio, and the blocks joined at its head and tail ("headjoined" /
 * "tailjoined" by index; none = freeEnd).  A block j named at an end of
 * block i is joined at j's tail if j's tail list names i, else at j's head
 * (PsBlockCopolymer::buildCntBlockType, pspolymer/PsBlockCopolymer.cpp:
 * 157-182).  Trees only: linear multiblocks, stars, combs. */
#define PSX_MAX_BLOCKS 16
#define PSX_MAX_JOIN 8
typedef struct psx_block {
  int species;
  double lengthfrac;
  double bSegRatio;
  int nhead, ntail;
  int head[PSX_MAX_JOIN], tail[PSX_MAX_JOIN];
} psx_block;

typedef struct psx_scft psx_scft;

/* Build: block steps (psbase/PsBlockBase.cpp:252-326, error
 * PSX_ERR_STEPS), exp(-ds b^2 k^2) tables (PsFlexPseudoSpec::build_k2,
 * pspolymer/PsFlexPseudoSpec.cpp:188-206), device buffers for all
 * contour slices.  Fields start at w = 0, p = 0. */
psx_status psx_scft_create(const psx_grid* grid, const psx_diblock* chain,
                           psx_prec prec, void* stream, psx_scft** out);
/* Same for a general block graph of A and B blocks sharing one ds
 * (PsBlockCopolymer with flexPseudoSpec blocks; volfrac = 1, the polymer's
 * length is the scale length).  Propagation follows PsBlockCopolymer::update
 * (pspolymer/PsBlockCopolymer.cpp:189-313): free ends start at q = 1, a
 * joined end at the product of the connected blocks' final propagators
 * (PsBlock::combineSetJnt, pspolymer/PsBlock.cpp:387-417); Q is block 0's
 * mean(qHeadInitial qHeadFinal).  PSX_ERR_STEPS for a non-integral block,
 * PSX_ERR_INVALID for a bad index, a repeated join, or a graph with a cycle
 * or an unreachable junction. */
psx_status psx_scft_create_blocks(const psx_grid* grid, int nblocks,
                                  const psx_block* blocks, double ds,
                                  double chiN, psx_prec prec, void* stream,
                                  psx_scft** out);
/* One polymer of a blend (a <Polymer> of kind blockCopolymer,
 * pspolymer/PsPolymerHldr.cpp:35-83): volfrac, length (the first polymer's
 * length is the scale length, lengthRatio = length / scale length), and its
 * blocks: the next nblocks entries of the block array, whose head / tail
 * indices count within the polymer. */
#define PSX_MAX_POLYMERS 8
typedef struct psx_polymer {
  double volfrac;
  double length;
  int nblocks;
} psx_polymer;

/* Same for a blend of npolymers block copolymers sharing the A and B fields
 * (e.g. a diblock with a homopolymer): block steps = nint(lengthRatio *
 * lengthfrac / ds) (psbase/PsBlockBase.cpp:192-197); each polymer has its
 * own Q from its first block; each block adds densWt * volfrac /
 * lengthRatio * qqt / Q_p to its species (pspolymer/PsBlock.cpp:170-173);
 * the species averages are sum volfrac * lengthfrac (psbase/PsBlockBase.cpp:
 * 185-189); F takes -sum_p volfrac ln Q_p / lengthRatio
 * (pseffhamil/PsPolymerUpdater.cpp:198-207).  chiN = chi * the first
 * polymer's length.  Contour checkpointing, dF/dfA and batches are
 * diblock / single-chain features (PSX_ERR_INVALID or PSX_ERR_STATE). */
psx_status psx_scft_create_blend(const psx_grid* grid, int npolymers,
                                 const psx_polymer* polymers,
                                 const psx_block* blocks, double ds,
                                 double chiN, psx_prec prec, void* stream,
                                 psx_scft** out);
psx_status psx_scft_destroy(psx_scft* e);
/* Block count, propagation paths (stored slice arrays) and rounds (sets of
 * paths that run concurrently; a path in round r starts from finals of
 * paths of earlier rounds). */
psx_status psx_scft_graph_info(const psx_scft* e, int* h_nblocks,
                               int* h_npaths, int* h_nrounds);
/* Slice s of block b's propagator from its head (from_tail = 0, s counted
 * from the head) or from its tail (s counted from the tail), FP64 out
 * (the reference's qs / qts of PsBlock, pspolymer/PsBlock.cpp:256-300). */
psx_status psx_scft_get_block_slice_host(psx_scft* e, int block,
                                         int from_tail, int s, double* h_out);
/* Switch the sweep precision (allocates the other slice set on first use).
 * The Anderson history is stored in the sweep's type: FP32 -> FP64 converts
 * it exactly and the history carries on (call psx_scft_reset_history for a
 * fresh one); FP64 -> FP32 restarts it.  Engines of a batch switch through
 * psx_batch_set_precision (PSX_ERR_STATE here). */
psx_status psx_scft_set_precision(psx_scft* e, psx_prec prec);
/* Contour steps of the A blocks and of the B blocks (for a diblock, the
 * block steps of A and B). */
psx_status psx_scft_block_steps(const psx_scft* e, int* sA, int* sB);

/* Conjugate fields w_A, w_B and pressure p, FP64 device arrays of N cells
 * (PsMonomerDens conjugate field, PsConstraint pressure field,
 *  psphysf/PsMonomerDens.h:93-132, psphysf/PsConstraint.h:95-104). */
psx_status psx_scft_set_fields(psx_scft* e, const double* d_wA,
                               const double* d_wB, const double* d_p);
psx_status psx_scft_get_fields(psx_scft* e, double* d_wA, double* d_wB,
                               double* d_p);
psx_status psx_scft_set_fields_host(psx_scft* e, const double* h_wA,
                                    const double* h_wB, const double* h_p);
psx_status psx_scft_get_fields_host(psx_scft* e, double* h_wA, double* h_wB,
                                    double* h_p);
/* The reference's host random stream theRng (PsRandom, psstd/PsRandom.h:
 * 147-170,305-373, psstd/PsRandom.cpp:58-95): MT19937 seeded by
 * init_genrand(seed mod 2^32) (PsDomain seeds it with randomSeed + rank,
 * pstrol/PsDomain.cpp:213); uniform = psRandom (y / 2^32), gaussian =
 * psRandomGaussian (Marsaglia polar, the second deviate kept for the next
 * call).  Host arrays; one stream may feed the initial fields and an
 * engine's noise, in the reference's order. */
typedef struct psx_rng psx_rng;
psx_status psx_rng_create(unsigned long long seed, psx_rng** out);
psx_status psx_rng_destroy(psx_rng* rng);
psx_status psx_rng_uniform(psx_rng* rng, long long n, double* h_out);
psx_status psx_rng_gaussian(psx_rng* rng, long long n, double* h_out);
/* Steepest-descent noise (PsPolymerUpdater::addRandUpdateConjgFields,
 * pseffhamil/PsPolymerUpdater.cpp:264-295, called from PsSteepDUpdater at
 * pseffhamil/PsSteepDUpdater.cpp:283): every steepest-descent step draws 2N
 * deviates from rng (w_A's cells, then w_B's) and adds strength * g to the
 * updated fields before the pressure and F.  rng NULL: no noise (and no
 * draws); strength 0 with an rng draws and adds nothing, as the reference.
 * The engine keeps the pointer: rng must outlive its use.  Anderson
 * mixing with noise on is PSX_ERR_INVALID. */
psx_status psx_scft_set_noise(psx_scft* e, double strength, psx_rng* rng);
/* Spectral filter updater (PsSimpleSpecFilter / PsMultiSpecFilter::update,
 * pseffhamil/PsSimpleSpecFilter.cpp:74-103) on the engine's conjugate
 * fields in place: fields is a mask of 1 (w_A) and 2 (w_B); the other
 * arguments as psx_fft_spec_filter (bulk: no wall volume).  The next
 * iteration propagates in the filtered fields. */
psx_status psx_scft_spec_filter(psx_scft* e, int fields, double cutoff_factor,
                                double filter_strength, const int* spec_cells);

/* w-factors and all block propagators (PsFlexPseudoSpec::reset + solveQ,
 * pspolymer/PsFlexPseudoSpec.cpp:159-181,244-293; block-graph order of
 * PsBlockCopolymer::update, pspolymer/PsBlockCopolymer.cpp:189-239). */
psx_status psx_scft_propagate(psx_scft* e);
/* Q (PsBlock::calcBigQ, pspolymer/PsBlock.cpp:205-237): PSX_ERR_NAN on NaN. */
psx_status psx_scft_Q(psx_scft* e, double* h_Q);
/* Q of every polymer (one per polymer of a blend, polymer order; a single
 * polymer gives psx_scft_Q's value); *h_n receives the polymer count. */
psx_status psx_scft_polymer_Q(psx_scft* e, double* h_Q, int* h_n);
/* Densities from the contour quadrature (PsBlock::setCalcQQTIntegral +
 * setPhysFields, pspolymer/PsBlock.cpp:154-196,256-380), FP64 device;
 * either output may be NULL (the engine keeps its own copy). */
psx_status psx_scft_densities(psx_scft* e, double* d_phiA, double* d_phiB);
/* Host copies of the densities of the last densities()/iterate() call. */
psx_status psx_scft_get_densities_host(psx_scft* e, double* h_phiA,
                                       double* h_phiB);
/* Copy slice t of propagation path `path` (FP64 out).  For a linear chain,
 * path 0 is the forward sweep (q of block 0 from its head, then on through
 * the chain) and path 1 the backward sweep; for the diblock: path 0 holds
 * q_A[0..sA] then q_B[1..sB], path 1 q+_B[0..sB] then q+_A[1..sA]. */
psx_status psx_scft_get_slice_host(psx_scft* e, int path, int t,
                                   double* h_out);

/* One steepest-descent update from the current densities
 * (PsSteepDUpdater::update, pseffhamil/PsSteepDUpdater.cpp:159-203; with
 * Flory calcDfD, PsConstraint::updatePres and calcFeField) -> F in the
 * reference order (after the w update). */
psx_status psx_scft_update_steepest(psx_scft* e, double lambda0,
                                    double lambda1, double* h_F);

/* Field-update method for psx_scft_iterate. */
typedef struct psx_mixing {
  int kind;              /* 0 steepest descent (lambda0, lambda1), 1 Anderson */
  double lambda0, lambda1;
  int depth;             /* Anderson history depth m (<= 32) */
  double beta;           /* Anderson mixing parameter */
  double anderson_start; /* Picard (lambda0) while residual RMS > this */
  int f_order;           /* 0 reference order (F after update), 1 consistent */
} psx_mixing;

/* Run up to n_iters full SCFT iterations (PsDomain::update,
 * pstrol/PsDomain.cpp:332-372), stopping early when the residual RMS
 * sqrt(mean(dH_A^2 + dH_B^2)) < tol (tol <= 0: never).  Histories (host,
 * length n_iters, may be NULL) receive F, Q and the residual RMS of each
 * iteration.  *h_done = iterations run.  use_graph != 0 captures one
 * iteration as a hipGraph and replays it. */
psx_status psx_scft_iterate(psx_scft* e, const psx_mixing* mix, int n_iters,
                            double tol, int use_graph, double* h_F,
                            double* h_Q, double* h_err, int* h_done);
/* Clear the Anderson history (e.g. after a precision switch). */
psx_status psx_scft_reset_history(psx_scft* e);

/* Free energy of the current fields and densities (reference order). */
psx_status psx_scft_free_energy(psx_scft* e, double* h_F);

/* dF/dL_x,y,z at fixed fields from the stored propagators (adjoint of the
 * split-step sweep); no reference counterpart (the reference has no
 * analytic stress).  Requires a preceding propagate. */
psx_status psx_scft_box_stress(psx_scft* e, double* h_dFdL);

/* Reverse mode of the forward op (w_A, w_B, box lengths) -> (phi_A, phi_B,
 * Q) at the current fields, for a scalar L = <g_phiA, phi_A> +
 * <g_phiB, phi_B> + g_Q Q (sums over cells): given the cotangents (device,
 * N each, NULL = zero) it writes dL/dw_A and dL/dw_B (device, N each) and,
 * when h_dLdbox is not NULL, dL/dL_x,y,z (host, 3).  The adjoint of each
 * stored propagator runs as one more sweep with the same step operators,
 * sources from the other propagator; the box terms reuse the stress's
 * spectral sums.  With g_phi = 0, g_Q = -1/Q it gives the derivatives of
 * -ln Q, the box stress at fixed fields.  Requires a preceding propagate,
 * FP64 sweeps with every slice stored, and one linear chain oriented head to
 * tail from block 0 (the diblock, linear multiblocks): PSX_ERR_STATE /
 * PSX_ERR_INVALID otherwise.  No reference counterpart. */
psx_status psx_scft_backward(psx_scft* e, const double* d_gphiA,
                             const double* d_gphiB, double g_Q, double* d_dLdwA,
                             double* d_dLdwB, double* h_dLdbox);
/* Same with host arrays. */
psx_status psx_scft_backward_host(psx_scft* e, const double* h_gphiA,
                                  const double* h_gphiB, double g_Q,
                                  double* h_dLdwA, double* h_dLdwB,
                                  double* h_dLdbox);

/* Contour quadrature of the densities: 0 = the reference's Simpson rule
 * with the literal 0.33333 and a trapezoid tail (pspolymer/PsBlock.cpp:
 * 256-380; the default), 1 = trapezoid weights ds [1/2, 1, ..., 1, 1/2],
 * the exact derivative of the split-step Q with respect to w, under which
 * the discrete free energy is stationary at the fixed point and the box
 * stress is its exact total derivative.  Mode 1 has no counterpart in the
 * reference. */
psx_status psx_scft_set_quadrature(psx_scft* e, int mode);
/* Which kernel sums the contour quadrature (no reference counterpart):
 * 0 (default) lets FP32 sweeps on grids of up to 64^3 cells (4-cell vectors
 * that fit one round of 64 per block) split each cell's contour sum over a
 * block's four waves; 1 always sums each cell serially in slice order.  The
 * two agree to rounding. */
psx_status psx_scft_set_quad_kernel(psx_scft* e, int kind);

/* Contour checkpointing for long chains (no reference counterpart): with
 * every = K > 0 the backward sweep keeps only its slices 0, K, 2K, ... and
 * its final slice.  Densities, the iteration and the box stress then
 * recompute K backward slices at a time from a checkpoint while the forward
 * propagator advances over the points that pair with them (one extra
 * backward sweep; memory per direction about Ns/K + 2K slices instead of
 * Ns + 1).  every = 0 stores every slice again (the default).  Linear chains
 * with every block oriented head to tail from block 0 only (PSX_ERR_INVALID
 * otherwise); engines of a batch cannot switch (PSX_ERR_STATE).  While on,
 * slice access, psx_scft_param_grad, psx_scft_time_step and
 * psx_scft_autotune return PSX_ERR_STATE.  Fields are kept; propagate again
 * after switching. */
psx_status psx_scft_set_checkpointing(psx_scft* e, int every);

/* Change the cell sizes (box lengths / grid counts) in place: rewrites the
 * exp(-ds b^2 k^2) tables (PsFlexPseudoSpec::build_k2 for the new
 * cellSizes); fields are kept.  Used by the box relaxation. */
psx_status psx_scft_set_cell(psx_scft* e, double dx, double dy, double dz);

/* Derivatives of the free energy along the fixed points of the field
 * update (evaluate at a converged state): dF/dchiN = mean(phiA phiB) - fA fB
 * and dF/dfA = -(1/Q) mean(q(fA) q+(fA) (wB - wA)), the derivative with
 * respect to the junction position (the discrete F moves in whole contour
 * steps, so finite differences agree to O(ds^2)).  Requires a preceding
 * propagate and densities.  dF/dfA is defined for the AB diblock only
 * (PSX_ERR_INVALID otherwise; pass NULL to get dF/dchiN of any graph).
 * No reference counterpart. */
psx_status psx_scft_param_grad(psx_scft* e, double* h_dFdchiN,
                               double* h_dFdfA);

/* Average device time (ms) of one contour step of the dual sweep and of
 * each of its line passes (z, y, x), timed with hipEvents on the engine's
 * stream over `reps` repetitions.  In plane mode (psx_scft_set_sweep_mode)
 * the first figure is the plane launches (z and y together), the second 0. */
psx_status psx_scft_time_step(psx_scft* e, int reps, double* h_ms_step,
                              double* h_ms_pass3);

/* Time the candidate line kernels of each pass (register splits, LDS
 * Stockham) on this engine's data and keep the fastest; h_choice (3 ints,
 * may be NULL) receives the chosen variant per pass (see
 * psx_scft_set_variant).  Where plane launches are available (and the
 * engine is not a batch's), whole sweeps in both sweep modes are timed
 * with the tuned line kernels and the faster mode is kept.  No reference
 * counterpart. */
psx_status psx_scft_autotune(psx_scft* e, int reps, int* h_choice);
/* Choose the line kernel of pass `pass` (grid axes with n > 1 from the
 * last, z, to the first, x) for the current precision: -1 the default for
 * the line length, 1 LDS Stockham, 2 and 3 the alternative register splits
 * (n = 64: 16x4; n = 96: 24x4, -; n = 128: 32x4 and 16x8; n = 32: -,
 * 4x8), 4 the 128-byte LDS row tiles for strided passes at n = 64, 96, 128
 * (the FP32 default there), 5 the
 * lane-pair register split 8x16 at n = 128 (the FP64 default); a variant
 * that does not apply to a pass uses the default.  No reference
 * counterpart. */
psx_status psx_scft_set_variant(psx_scft* e, int pass, int variant);
/* How a sweep is launched (no reference counterpart; results agree to
 * rounding).  mode 0: line launches, two per contour step (a strided pass
 * alone, then a fused pass that ends one step and starts the next).  mode 1:
 * plane launches, one per contour step -- each x-plane's z and y round
 * trips of two consecutive steps run in one launch that stages the plane in
 * LDS, alternating with the fused x launches, so a step moves 4 fields
 * through HBM instead of 6.  Mode 1 needs a 3D grid with ny == nz in
 * {32, 48, 64, 96, 128} (PSX_ERR_INVALID otherwise) and is the default there.
 * Switching rebuilds the sweep buffers: fields are kept, propagate again.
 * Engines of a batch switch through psx_batch_set_sweep_mode
 * (PSX_ERR_STATE here).  The environment variable PSX_PLANE=0 makes line
 * launches the default of new engines. */
psx_status psx_scft_set_sweep_mode(psx_scft* e, int mode);
/* Current sweep mode and whether plane launches are available (either
 * pointer may be NULL). */
psx_status psx_scft_sweep_mode(const psx_scft* e, int* h_mode, int* h_available);

/* ===================================================================== */
/* Batched parameter sweeps: many independent diblock instances (grid,     */
/* box, fA, chiN may all differ; the contour step count must match)        */
/* advanced together.  Instances with the same grid share each line-pass   */
/* launch (both sweeps of every member), the per-instance reductions and   */
/* updates follow, and one batch iteration is one hipGraph per group of at */
/* most four size classes (groups run one after another).  No reference   */
/* counterpart (the reference runs one system per process).                */
/* ===================================================================== */
typedef struct psx_batch psx_batch;
psx_status psx_batch_create(int n, const psx_grid* grids,
                            const psx_diblock* chains, psx_prec prec,
                            void* stream, psx_batch** out);
psx_status psx_batch_destroy(psx_batch* b);
/* Instance count and number of distinct grid sizes. */
psx_status psx_batch_size(const psx_batch* b, int* n, int* nclasses);
/* The engine of instance i (owned by the batch): fields, densities, stress
 * and the other psx_scft_* calls work on it directly, except the switches
 * that change which buffers the class launches write (precision, sweep mode,
 * checkpointing, noise, spectral filters: PSX_ERR_STATE). */
psx_status psx_batch_instance(psx_batch* b, int i, psx_scft** out);
psx_status psx_batch_set_precision(psx_batch* b, psx_prec prec);
/* Per grid size and pass, time the candidate line kernels on the batch's
 * own fields and keep the fastest; h_choice (3 per size class, may be
 * NULL) receives the chosen variants (-1 register default, 1 Stockham,
 * 2 alternative register split).  Classes where plane launches are
 * available then time whole class sweeps in both sweep modes and keep the
 * faster. */
psx_status psx_batch_autotune(psx_batch* b, int reps, int* h_choice);
/* Sweep mode of size class cls (see psx_scft_set_sweep_mode; class order
 * is that of first appearance among the instances).  Planes are the
 * default for classes of n < 96 and for classes of at most four
 * instances. */
psx_status psx_batch_set_sweep_mode(psx_batch* b, int cls, int mode);
psx_status psx_batch_sweep_mode(const psx_batch* b, int cls, int* h_mode);
/* n_iters iterations of every instance; histories are n * n_iters, row-major
 * by instance (may be NULL). */
psx_status psx_batch_iterate(psx_batch* b, const psx_mixing* mix, int n_iters,
                             int use_graph, double* h_F, double* h_err,
                             int* h_done);

#ifdef __cplusplus
}
#endif
#endif
