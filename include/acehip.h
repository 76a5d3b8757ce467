/*
 * acehip.h — C ABI of libacehip.so, the MI355X (gfx950) basis-evaluation
 * path of the Atomic Cluster Expansion with ACE.jl (v0.12.46) semantics.
 *
 * The library replaces the evaluator slot of ACE.jl's LinearACEModel
 * (src/linearmodel.jl:44-56: ``evaluator = :standard`` -> ProductEvaluator,
 * src/evaluator.jl:9-13) and the basis-level evaluate / evaluate_d of
 * Product1pBasis, PIBasis and SymmetricBasis.  A host (Julia ``ccall``,
 * Python ``ctypes``) builds the basis on the CPU, hands the flat tables to
 * ``ace_ctx_create`` and then calls the batched entry points below.
 *
 * Conventions
 *  - Return value 0 = success, < 0 = error; ``ace_last_error()`` gives the
 *    message (thread-local).  Errors map to the reference's exceptions.
 *  - All FP64.  Complex values are interleaved (re, im).
 *  - Environments are ragged: ``env_offsets[nenv+1]`` (int32, ascending,
 *    env_offsets[0] == 0) index neighbour arrays.  Neighbour positions are
 *    SoA ``xyz[3][nnbr]`` (x block, then y, then z), relative to the centre.
 *  - Every array argument of an evaluation call is a DEVICE pointer
 *    (HBM-resident, e.g. a torch or hipMalloc buffer); ``stream`` is a
 *    hipStream_t (NULL = default stream).  Calls are stream-ordered and
 *    return without synchronizing.  The library keeps no caller pointer past
 *    the call.  One context must not be used by two calls at the same time.
 *  - ``ace_ctx_create`` / ``ace_set_params`` take HOST pointers.
 */
#ifndef ACEHIP_H
#define ACEHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACEHIP_ABI_VERSION 3

typedef struct ace_ctx ace_ctx;

#define ACE_RUN_W 11      /* int32 words per run of the device programs */
#define ACE_MAX_ORDER 8   /* largest correlation order of a basis */

/* Property of the symmetric basis: Invariant (properties.jl:77-152) or
 * EuclideanVector (properties.jl:165-248). */
enum { ACE_PROP_INVARIANT = 0, ACE_PROP_EUCLIDEAN_VECTOR = 1 };

/*
 * Flat basis definition (host memory; copied by ace_ctx_create).
 *
 * Radial basis (orthpolys.jl:82-99, distancetransforms.jl:73):
 *   t = ((1 + r0) / (1 + r))^tp, P1 = rA[0] fcut(t),
 *   P2 = (rA[1] t + rB[1]) P1, Pn = (rA[n] t + rB[n]) P(n-1) + rC[n] P(n-2),
 *   fcut = (t - tl)^pl (t - tr)^pr, zero for t < tl (pl > 0) / t > tr (pr > 0).
 * Pooled 1p table: pool[k] = (mu, n, l, m) with m >= 0, n 1-based; the 1p
 *   value A+[k] = sum_j [mu_j == mu] R_n(r_j) Y_l^m(rhat_j).  mu = -1 when
 *   the 1p basis has no categorical component.
 * 1p table of the basis: aspec_fid[iA] = k | 0x8000 * (m < 0) and
 *   aspec_sgn[iA] = (-1)^m for m < 0 (Y_l^-m = (-1)^m conj Y_l^m).
 * AA table (PIBasisSpec, pibasis.jl:10-13): orders[nAA], iAA2iA[nAA][maxorder]
 *   (0-based A indices, descending, -1 padding).
 * A2B map (SymmetricBasis.A2Bmap) in CSR: a2b_rowptr[nB+1], a2b_cols[nnz],
 *   a2b_vals[nnz * (1 | 6)] (real, or 3 complex for EuclideanVector).
 * Programs: run programs for P = sum c~ AA and its gradient polynomials
 *   G[k] (compiled by ace_ctx_create_jl from the basis tables, or by the
 *   Python host); leaf coefficients are linear in c~: leaf_terms[leaf] =
 *   (term_begin) into term arrays (term_row, term_scale, term_conj);
 *   task_terms likewise for task constants.
 */
typedef struct ace_pack {
    int32_t abi_version;
    int32_t prop;
    int32_t L;
    int32_t nr;
    int32_t nspecies;
    int32_t tp;
    double r0, tl, tr;
    int32_t pl, pr;
    const double* rA;
    const double* rB;
    const double* rC;
    int32_t npool;
    const int32_t* pool;          /* [npool][4] */
    int32_t nA;
    const int32_t* aspec_fid;     /* [nA] */
    const double* aspec_sgn;      /* [nA] */
    int32_t nAA, maxorder;
    const int32_t* orders;        /* [nAA] */
    const int32_t* iAA2iA;        /* [nAA][maxorder] */
    int32_t nB, a2b_nnz;
    const int32_t* a2b_rowptr;    /* [nB+1] */
    const int32_t* a2b_cols;      /* [nnz] */
    const double* a2b_vals;       /* [nnz * ncoef] */
    /* run programs */
    int32_t nruns;
    const int32_t* runs;          /* [nruns][ACE_RUN_W] depth, jstart, nleaf, leaf_off, prefix
                                     factors f1..f7 (depth <= 7: correlation order <= 8) */
    int32_t nleaves;
    const int32_t* leaf_fid;      /* [nleaves] */
    const int32_t* leaf_term_ptr; /* [nleaves+1] */
    int32_t ntasks;
    const int32_t* tasks;         /* [ntasks][5] run_begin, run_end, out_slot (-1 energy),
                                     rec_begin, rec_end */
    const int32_t* task_term_ptr; /* [ntasks+1] constant terms */
    int32_t nterms;
    const int32_t* term_row;      /* [nterms] AA row of c~ */
    const double* term_scale;     /* [nterms] */
    const int32_t* term_conj;     /* [nterms] 1: conj(c~ * scale) */
    int32_t ngslots;              /* number of G output slots */
    const int32_t* gslot_ptr;     /* [npool+1] slots summed into G[k] */
    const int32_t* gslot_list;    /* [ngslots] */
    /* EuclideanVector programs split by class (1) or by component (0): with
     * classed = 1 every AA column's A2B values are one complex number times
     * a fixed unit vector u_M of its class M = total m (in {-1, 0, 1}), the
     * energy tasks of class c = M + 1 have output slot -1 - c, and
     * gslot_class[slot] is the class of a G slot. */
    int32_t classed;
    const int32_t* gslot_class;   /* [ngslots] when classed */
    /* the same programs as 64-byte records (Invariant, correlation order <= 4;
     * nrec = 0 otherwise): 8 header dwords (dw0 = depth | jstart << 2 |
     * nleaf << 5, dw1 = f1 | f2 << 16, dw2 = f3, dw3.. leaf factor ids two per
     * dword) + 4 coefficients filled from rec_leaf */
    int32_t nrec;
    const uint32_t* rec_hdr;      /* [nrec][8] */
    const int32_t* rec_leaf;      /* [nrec][4] leaf index per slot, -1 empty */
    /* order <= 2 invariants (surrogate inputs): feature f = sum over terms
     * z in [feat_ptr[f], feat_ptr[f+1]) of feat_coef[z] * Re(A_f1 A_f2) with
     * (f1, f2) = feat_f[z][0..1] device factor ids (-1 = absent); feat_rows[f]
     * is its B row.  Gradient entries per pooled index k: fgrad_ptr[k]..,
     * (feature, other factor, coef): G[k] += w_feature coef value(other). */
    int32_t nfeat;
    const int32_t* feat_rows;     /* [nfeat] */
    const int32_t* feat_ptr;      /* [nfeat+1] */
    const int32_t* feat_f;        /* [nterm][2] */
    const double* feat_coef;      /* [nterm] */
    const int32_t* fgrad_ptr;     /* [npool+1] */
    const int32_t* fgrad_feat;    /* [ngrad] */
    const int32_t* fgrad_f;       /* [ngrad] */
    const double* fgrad_coef;     /* [ngrad] */
    /* task -> wave assignment for each supported waves-per-tile setting */
    int32_t nwave_cfg;
    const int32_t* wave_cfg_W;    /* [nwave_cfg] waves per workgroup */
    const int32_t* wave_task_ptr; /* concat over cfgs: [W+1] each (energy+grad) */
    const int32_t* wave_task_list;/* concat over cfgs: task ids */
    const int32_t* wave_etask_ptr;/* concat over cfgs: [W+1] each (energy only) */
    const int32_t* wave_etask_list;
} ace_pack;

/* Create a context on HIP device ``device`` with coefficients c[nB]
 * (LinearACEModel(basis, c), linearmodel.jl:44-56; c~ = A2B^T c,
 * evaluator.jl:30-31,59-67). */
int ace_ctx_create(const ace_pack* pack, int device, const double* c, int64_t nB,
                   ace_ctx** out);

/*
 * A basis as ACE.jl holds it: the fields of SymmetricBasis, PIBasis,
 * Product1pBasis and the radial OrthPolyBasis in their Julia memory layout
 * (Int = int64, 1-based indices, column-major Matrix, SparseMatrixCSC).  A
 * Julia caller passes pointers to the objects' own arrays; nothing derived.
 *
 *  - radial basis: OrthPolyBasis fields pl, tl, pr, tr, A, B, C
 *    (orthpolys.jl:82-99) and the transform t = ((1 + r0)/(1 + r))^tp
 *    (PolyTransform, distancetransforms.jl:16,73);
 * *  - maxL: the degree bound of the SHBasis (Ylm component), at most 10;
 *  - ncategories: length of the Categorical1pBasis (0 when absent);
 *  - Product1pBasis.indices (product_1pbasis.jl:5-8), Vector{NTuple{ncomp,
 *    Int}} -> indices[nA][ncomp], with comp_kind[c] the kind of bases[c]:
 *    ACE_COMP_CATEGORICAL (category index), ACE_COMP_RN (n), ACE_COMP_YLM
 *    (index_y(l, m) = m + l + l^2 + 1, sphericalharmonics.jl:102);
 *  - PIBasis.spec (pibasis.jl:10-13): orders[nAA], iAA2iA column-major
 *    [maxorder][nAA] (entry (i, t) at iAA2iA[(t - 1) * nAA + i - 1]), 0 padding;
 *  - SymmetricBasis.A2Bmap (symmbasis.jl:33-38), nB x nAA CSC: colptr[nAA+1],
 *    rowval[nnz], nzval[nnz] (Invariant{Float64}) or nzval[nnz][6]
 *    (EuclideanVector{ComplexF64}: 3 complex components, re/im interleaved).
 */
enum { ACE_COMP_CATEGORICAL = 0, ACE_COMP_RN = 1, ACE_COMP_YLM = 2 };

typedef struct ace_jl_basis {
    int32_t abi_version;
    int32_t prop;                 /* ACE_PROP_INVARIANT / ACE_PROP_EUCLIDEAN_VECTOR */
    int64_t nr;
    const double* A;
    const double* B;
    const double* C;
    int64_t pl, pr;
    double tl, tr;
    int64_t tp;
    double r0;
    int64_t maxL;
    int64_t ncategories;
    int64_t ncomp;
    const int32_t* comp_kind;     /* [ncomp] */
    int64_t nA;
    const int64_t* indices;       /* [nA][ncomp] */
    int64_t nAA, maxorder;
    const int64_t* orders;        /* [nAA] */
    const int64_t* iAA2iA;        /* [maxorder][nAA] */
    int64_t nB;
    const int64_t* colptr;        /* [nAA + 1] */
    const int64_t* rowval;        /* [nnz] */
    const double* nzval;          /* [nnz] or [nnz][6] */
} ace_jl_basis;

/* Context from ACE.jl's own tables: ProductEvaluator(basis, c)
 * (evaluator.jl:30-31 via LinearACEModel(basis, c; evaluator),
 * linearmodel.jl:44-56).  Compiles the device programs (pooled 1p table,
 * product-DAG runs and records, adjoint polynomials, task lists) on the host
 * and creates the context as ace_ctx_create does. */
int ace_ctx_create_jl(const ace_jl_basis* basis, int device, const double* c, int64_t nB, ace_ctx** out);

/* Host-only: compiles the device tables of ``basis`` without a device and
 * reports their sizes (npool, nruns, nleaves, ntasks, nrec, ngslots, nfeat
 * in sizes[0..6]).  Returns the same errors as ace_ctx_create_jl. */
int ace_compile_check(const ace_jl_basis* basis, int64_t* sizes);

/* set_params!(model, c) (linearmodel.jl:67-71 -> evaluator.jl:48-57). */
int ace_set_params(ace_ctx* ctx, const double* c, int64_t nB);

/* Tile shape: waves per workgroup of the product-DAG kernel (must be one of
 * the pack's wave configurations). */
int ace_set_tile(ace_ctx* ctx, int waves_per_tile);

/* Workgroups per product-DAG tile (Invariant): each runs 1/parts of the
 * tile's tasks, which shortens the last, partly empty round of workgroups.
 * 0 (default) chooses 1 to 4 per call from the tile and CU counts. */
int ace_set_parts(ace_ctx* ctx, int parts);

/* Tiles at the end of the product-DAG grid (record kernel) that run one more
 * part than the rest, so that the last round holds shorter workgroups.
 * -1 (default) chooses the count per call when the parts are automatic;
 * n >= 0 uses n tiles and needs 2 or 3 parts per tile.  A tuning knob of
 * this implementation, like ace_set_parts; no counterpart in the reference. */
int ace_set_tail_parts(ace_ctx* ctx, int ntail);

/* The product-DAG launch of the last evaluation call: workgroups per tile,
 * the number of tail tiles that ran one more part, and the kernel (0: the
 * record kernel on 64-environment tiles, 1: on 32-environment tiles; 64, 32
 * or 16: the run interpreter with that many environments per tile).
 * Instrumentation; no counterpart in the reference. */
int ace_last_dag_shape(ace_ctx* ctx, int* parts, int64_t* tail_tiles, int* kernel);

/* Product-DAG kernel choice: 0 (default) the record-segment kernel where the
 * basis allows it (Invariant or class-split EuclideanVector programs,
 * correlation order <= 4, at most 320 pooled values), else the run
 * interpreter; 1 the run interpreter always; 2 the per-record dispatch kernel
 * over the same records where they exist.  A tuning and testing knob of this
 * implementation; no counterpart in the reference. */
int ace_set_kernel(ace_ctx* ctx, int mode);

/* Pooling launch shape: neighbours staged per chunk (8, 16 or 32; 0 = 16)
 * and environments per wave (0 = automatic: two for bases of at most 32
 * pooled values, else one; 1 = one).  Results do not depend on it.  A tuning
 * knob of this implementation; no counterpart in the reference. */
int ace_set_pool_shape(ace_ctx* ctx, int chunk, int envs_per_wave);

/* Site energies, ProductEvaluator evaluate (evaluator.jl:121-147), batched.
 * E_out[nenv] (Invariant) or E_out[nenv][3] (EuclideanVector).  species
 * [nnbr] int8 category index (NULL without a categorical 1p basis). */
int ace_energy(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets,
               const double* xyz, const int8_t* species, double* E_out,
               void* stream);

/* Energies and grad_config (evaluator.jl:150-200): dE_dr[3][nnbr] SoA is
 * +dE/dr_j of the environment that owns neighbour j (ACE.jl sign);
 * EuclideanVector: dE_dr[9][nnbr], entry (p*3+d) = d phi_p / d r_j,d. */
int ace_energy_forces(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets,
                      const double* xyz, const int8_t* species, double* E_out,
                      double* dE_dr, void* stream);

/* Energies and gradients as ace_energy_forces, allocating nothing: no
 * device or host-pinned allocation and no host read of the inputs, so the
 * call can run under HIP graph capture or in a loop that must not allocate.
 * No counterpart in the reference; for callers that cannot allocate.  The
 * caller passes nnbr = env_offsets[nenv] (the offsets are not validated).
 * The scratch (pooled values, adjoint, energy shares, neighbour map) lives
 * in a fixed 64 MiB device array, one per device for the life of the
 * process: a batch that needs more is rejected with an error (at the
 * order-4 degree-12 basis about 12,000 environments of 30 neighbours fit).
 * Every call on one device shares that array, so calls that overlap on one
 * device (other contexts or streams included) race. */
int ace_energy_forces_noalloc(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, int64_t nnbr,
                              const double* xyz, const int8_t* species, double* E_out, double* dE_dr,
                              void* stream);

/* Host-memory variants for callers without device buffers (a Julia ccall
 * with Base arrays): every array argument is a HOST pointer with the layout
 * of the device entry point of the same name; env_offsets and species are
 * checked on the host, the inputs are staged to the context's device
 * buffers on the context's own stream, and the call returns after the
 * outputs are back (PCIe-bound: about 50 bytes per neighbour). */
int ace_energy_host(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, const double* xyz,
                    const int8_t* species, double* E_out);
int ace_energy_forces_host(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, const double* xyz,
                           const int8_t* species, double* E_out, double* dE_dr);
int ace_basis_B_host(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, const double* xyz,
                     const int8_t* species, double* B_out);
int ace_basis_dB_host(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, const double* xyz,
                      const int8_t* species, double* dB_out);
int ace_adjoint_EVAL_D_host(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, const double* xyz,
                            const int8_t* species, const double* w, double* dB_out);

/* Product1pBasis evaluate (product_1pbasis.jl:123-134): A[nenv][nA] complex. */
int ace_basis_A(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets,
                const double* xyz, const int8_t* species, double* A_out,
                void* stream);

/* PIBasis evaluate (pibasis.jl:265-275): AA[nenv][nAA], real (Invariant) or
 * complex (EuclideanVector). */
int ace_basis_AA(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets,
                 const double* xyz, const int8_t* species, double* AA_out,
                 void* stream);

/* SymmetricBasis evaluate (symmbasis.jl:297-315): B[nenv][nB] (Invariant)
 * or B[nenv][nB][3] (EuclideanVector). */
int ace_basis_B(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets,
                const double* xyz, const int8_t* species, double* B_out,
                void* stream);

/* SymmetricBasis evaluate_d (symmbasis.jl:322-336): dB[nnbr][nB][3]
 * (Invariant) or dB[nnbr][nB][3][3] (EuclideanVector, [out][space]). */
int ace_basis_dB(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets,
                 const double* xyz, const int8_t* species, double* dB_out,
                 void* stream);

/* adjoint_EVAL_D(m, ::ProductEvaluator, cfg, w) (evaluator.jl:204-244),
 * batched over environments (Invariant bases): with per-neighbour weights
 * w[3][nnbr] (SoA, device), dB_out[nenv][nB] = sum_j w_j . dB/dr_j, the
 * derivative of the basis along w without forming dB (linearmodel.jl:133-134;
 * the mixed second derivative behind force-matching gradients,
 * linearmodel.jl:195-229). */
int ace_adjoint_EVAL_D(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, const double* xyz,
                       const int8_t* species, const double* w, double* dB_out, void* stream);

/* Learned-surrogate support (no reference counterpart; for the surrogate
 * path of a LinearACEModel, see acemlp.h).
 *
 * ace_surr_features: pooled 1p values (kept in the context for the next
 * ace_surr_energy_forces call), the order-1 and order-2 invariants X[nfeat][nenv]
 * (FP64, nullable), their normalized BF16 copy Xn[nenv][160] =
 * (X - xmean) * xinvstd zero-padded (nullable), and the exact order <= 2
 * energy share E2[nenv] = c0 + sum_f c_f X_f (c0: the order-0 constant).  xmean / xinvstd are device
 * arrays [nfeat]. */
int ace_surr_features(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, const double* xyz,
                      const int8_t* species, const double* xmean, const double* xinvstd, double* X,
                      void* Xn, double* E2, void* stream);

/* Surrogate energies (and forces): with member predictions y[nmemb][nenv]
 * of the normalized order-3/4 share and (for forces) their input gradients
 * dX[nmemb][nenv][160], E = E2 + ymean + ystd * mean_m y_m.  Environments
 * whose ensemble spread ystd * std_m y_m exceeds tau fall back to plain ACE
 * (the product DAG on this context); their indices are written to
 * fallback_list and their count to *fallback_count (device int32).  Forces
 * of accepted environments follow from the chain rule through the features;
 * dE_dr as in ace_energy_forces.  Must follow ace_surr_features on the same
 * batch and context. */
int ace_surr_energy_forces(ace_ctx* ctx, int64_t nenv, const int32_t* env_offsets, const double* xyz,
                           const int8_t* species, int nmemb, const float* y, const float* dX,
                           double ymean, double ystd, const double* xinvstd, double tau,
                           const double* E2, double* E_out, double* dE_dr, int32_t* fallback_count,
                           int32_t* fallback_list, void* stream);

/* Kernel-time instrumentation: average duration in ms of the last call's
 * product-DAG kernel (measured with HIP events on the call's stream) and the
 * executed leaf/run counts per environment. */
int ace_last_timing(ace_ctx* ctx, double* dag_ms, double* pool_ms, double* force_ms);

/* Host-side validation of env_offsets (non-decreasing, starting at 0) and
 * species (known categories) before every evaluation call: on (default) or
 * off for callers that validated their buffers once. */
int ace_set_validation(ace_ctx* ctx, int on);

const char* ace_last_error(void);
void ace_ctx_destroy(ace_ctx* ctx);
int ace_abi_version(void);

/*
 * Environment extraction and force assembly: the atoms-level layer around the
 * evaluator.  ACE.jl evaluates one environment at a time; its callers build
 * the environments of a configuration and sum site energies and forces
 * (energy(V, at) / forces(V, at) in src/testing/testlsq.jl:27-35,42-69 and
 * src/testing/testmodel.jl:64-81; the neighbour list itself comes from the
 * JuLIP package, which ACE.jl v0.12.46 does not depend on).  Entry points
 * below replace that layer for whole configurations on the device.
 *
 * Positions are SoA pos[3][natoms] (device).  cell[9] (host) holds the
 * lattice vectors as rows, pbc[3] (host) 1 = periodic along that vector (the
 * cell must be non-singular also along open axes: there it only sets the
 * direction of the binning).  Entry m of atom i is the neighbour j at
 * r_m = (x_j - x_i) + S'_m . cell, 0 < |r_m| < rcut (self pairs excluded,
 * self images included), with integer image shift S'_m relative to the
 * given positions.  The list is exactly symmetric: (i -> j, S') present iff
 * (j -> i, -S') present, with the exactly negated vector.  The order within an
 * environment is deterministic.
 */
#define ACE_ERR_CAPACITY (-2)

typedef struct ace_nbrlist ace_nbrlist;

int ace_nbrlist_create(int device, ace_nbrlist** out);
void ace_nbrlist_destroy(ace_nbrlist* nl);

/* Builds the neighbour lists: env_offsets[natoms+1], and with
 * nnbr = env_offsets[natoms] entries, nbr_j[nnbr], shift[3][nnbr] (SoA, int32),
 * xyz[3][nnbr] (the evaluator's neighbour positions, stride nnbr), rev[nnbr]
 * (entry index of the reverse pair; nullable) and nbr_species[nnbr] =
 * atom_species[nbr_j] (nullable; int8 category index).  The arrays must hold
 * ``capacity`` entries (xyz and shift 3 * capacity).  Returns
 * ACE_ERR_CAPACITY with *nnbr_out = the entries needed when capacity is too
 * small (nothing else written).  Synchronizes ``stream``: the binning runs on
 * the host and the list length is returned on the host. */
int ace_nbrlist_build(ace_nbrlist* nl, int64_t natoms, const double* pos, const double* cell,
                      const int32_t* pbc, double rcut, const int8_t* atom_species, int64_t capacity,
                      int32_t* env_offsets, int32_t* nbr_j, int32_t* shift, double* xyz, int32_t* rev,
                      int8_t* nbr_species, int64_t* nnbr_out, void* stream);

/* Recomputes xyz[3][nnbr] = (x_j - x_i) + S' . cell for a list from
 * ace_nbrlist_build and new positions (and cell, passed by value: a captured
 * call keeps the cell it was captured with).  Stream-ordered, allocates
 * nothing, capturable in a HIP graph. */
int ace_nbrlist_update(ace_nbrlist* nl, int64_t natoms, const double* pos, const double* cell,
                       int64_t nnbr, const int32_t* env_offsets, const int32_t* nbr_j,
                       const int32_t* shift, double* xyz, void* stream);

/* Atomic forces F[3][natoms] = -dE/dx (E = sum of site energies) from the
 * per-neighbour gradients dE_dr[3][nnbr] of ace_energy_forces (same list),
 * and optionally the virial[9] = -sum_m dE_dr_m r_m^T (row-major,
 * device; needs xyz).  Deterministic (no atomics).  Stream-ordered; with a
 * virial the first call per size allocates a small partials buffer, later
 * calls of equal or smaller size allocate nothing. */
int ace_nbrlist_forces(ace_nbrlist* nl, int64_t natoms, const int32_t* env_offsets, const int32_t* rev,
                       int64_t nnbr, const double* xyz, const double* dE_dr, double* F, double* virial,
                       void* stream);

/* The same assembly for nvec fields at once: g[nnbr][nvec][3] (e.g. dB of
 * ace_basis_dB, one field per basis function) -> F[natoms][nvec][3] =
 * -d/dx of the summed field; the per-basis-function forces
 * ``forces(basis, at)`` of least-squares fitting (src/testing/testlsq.jl:
 * 55-61).  Stream-ordered, allocates nothing, deterministic. */
int ace_nbrlist_forces_multi(ace_nbrlist* nl, int64_t natoms, const int32_t* env_offsets,
                             const int32_t* rev, int64_t nnbr, int64_t nvec, const double* g, double* F,
                             void* stream);

#ifdef __cplusplus
}
#endif
#endif
