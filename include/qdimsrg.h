/*
 * qdimsrg.h -- C ABI of the MI355X (gfx950) IM-SRG(2) commutator / BCH engine.
 *
 * Plain C, no GPU or framework types in any signature.  Every entry point
 * returns QD_OK (0) or a negative QD_E* code; qd_last_error() gives the text
 * of the most recent failure on the calling thread.
 *
 * Data model (the reference's operator API, src/ME_general.f90:26-62):
 *   qd_context  one channel structure: N electrons in R oscillator shells,
 *               M = R(R+1) spin-orbitals, (M_L, M_S) two-body blocks and the
 *               cross-coupled (Pandya) blocks; owns one HIP stream and the
 *               device workspaces.  Replaces reindex / build_cross_coupled
 *               (src/ME_general.f90:956-1116, 363-448).
 *   qd_op       one device-resident type(full_ham): neq doubles in the
 *               reference's vectorize order (src/ME_general.f90:1871-1989) --
 *               E0, fhh, fpp, fph, then per block Vhhhh, Vpppp, Vpphh, Vphph,
 *               Vppph, Vphhh, column-major -- plus the integer herm flag.
 *   qd_cc       one device-resident type(cc_mat): per CC block an r x nb
 *               column-major matrix, plus the real herm flag
 *               (src/ME_general.f90:13-24).
 * Host vectors passed to upload/download are exactly the reference's
 * vectorize/repackage vectors, so the ODE state of src/adams_ode.f90 and a
 * Fortran full_ham round-trip through them unchanged.
 *
 * All work of one context runs in order on its stream; entry points that
 * return a host scalar synchronise that stream.  A context is not reentrant:
 * calls on one context from several threads must be serialised by the caller
 * (the reference's callers are single-threaded).
 */
#ifndef QDIMSRG_H
#define QDIMSRG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QD_OK 0
#define QD_E_ARG (-1)      /* bad argument (null handle, size mismatch, bad herm) */
#define QD_E_HIP (-2)      /* HIP runtime error (allocation, launch, copy)       */
#define QD_E_LAYOUT (-3)   /* channel structure not supported                    */
#define QD_E_NUMERIC (-4)  /* non-finite value met where the reference would      */
                           /* silently propagate it (flow drivers only)           */

typedef struct qd_context qd_context;
typedef struct qd_op qd_op;
typedef struct qd_cc qd_cc;

typedef struct {
  int64_t nbody, nshell, msp, nblock, ncc;
  int64_t neq;      /* doubles per operator (get_num_elements, src/ME_general.f90:1755) */
  int64_t cc_size;  /* doubles per CC matrix set (sum of r*nb)                          */
  int64_t n2b;      /* two-body doubles per operator                                    */
} qd_info;

const char *qd_last_error(void);

/* ---- context: channel structure + device tables -------------------------
 * Replaces arrange_states / reindex / build_cross_coupled
 * (src/ME_general.f90:688-744, 956-1116, 363-448).  Holes are the first
 * nbody states in states order (the reference's energy->states map is the
 * identity, convertF src/ME_general.f90:1527-1536); nbody must fill shells.
 * Environment read at creation (tuning only; results do not change):
 * QD_BOX_CAP grid elements per two-body box task (16..3600, default 3600),
 * QD_GATHER_TABLES=0 per-element calc_cc and 221 paths instead of the gather
 * tables, QD_GEMM_WG_PER_CU work-groups per CU of the persistent GEMM grid
 * (1..16, default 8), QD_BOX_PAIRS=0 every class-pair box of the two-body
 * pass evaluated on its own instead of paired with its transpose (results
 * then differ at rounding level only), QD_BOX_WIDE=0/1 the narrow (512
 * threads, 3600-element grids) or wide (1024 threads, 8000) box tasks instead
 * of the automatic choice (wide where the narrow form would split more than a
 * tenth of the grid elements into slabs; QD_BOX_CAP then ranges up to 8000),
 * QD_DSTAGE=0/1 whole boxes read their Pandya sub-blocks of D from HBM or
 * stage them in LDS (default: staged in the wide form), QD_ZSKIP=0 the BCH
 * terms run every GEMM product instead of skipping those that read a
 * sub-block type of G that is zero in every block (the same values),
 * QD_LDIRECT=0 the square two-body sub-blocks' ladder terms go through a
 * buffer filled by their own pass instead of being formed in the box pass
 * (the same values to rounding). */
int qd_context_create(int nbody, int nshell, qd_context **out);
int qd_context_destroy(qd_context *ctx);
int qd_context_info(const qd_context *ctx, qd_info *out);
/* Export one host table for inspection: which = 0 block dims (nblock x 3:
 * nhh,npp,nph), 1 flat offsets (nblock x 6), 2 i_array (M(M-1)/2 x 3: type
 * 1..3, block, index; 0-based), 3 CC map (M*M x 2: block,row), 4 CC dims
 * (ncc x 2: r, nb), 5 pair lists (hh, ph, pp of every block, concatenated,
 * x 2: (i, j) with i < j), 6 sizes (neq, cc_size, nblock, ncc), 7 single-
 * particle quantum numbers (M x 2: m_l, m_s as 2*s_z, states order), 8 the
 * calc_cc gather (cc_size x 2: per CC element in device order its code,
 * flat V index | stored transposed << 30 | pair-sorting sign << 31 or
 * 0x3fffffff for zero, and its flat index in the reference's CC order).
 * len is the capacity of out in int64 elements; returns the count. */
int64_t qd_context_table(const qd_context *ctx, int which, int64_t *out, int64_t len);
/* The same tables without a context or a GPU (host-only channel builder);
 * which = 6 gives (neq, cc_size, nblock, ncc). */
int64_t qd_layout_table(int nbody, int nshell, int which, int64_t *out, int64_t len);
/* Block until all work queued on the context's stream is done. */
int qd_context_sync(qd_context *ctx);

/* ---- operators: type(full_ham), allocate_everything --------------------
 * src/IMSRG_tools.f90:465-537 (allocate_everything: zeroed, herm = +1). */
int qd_op_create(qd_context *ctx, int herm, qd_op **out);
int qd_op_destroy(qd_op *op);
int qd_op_set_herm(qd_op *op, int herm);  /* full_ham%herm, +1 or -1 */
int qd_op_get_herm(const qd_op *op, int *herm);
/* repackage (src/ME_general.f90:1992-2109): host vectorize-order -> device */
int qd_op_upload(qd_op *op, const double *vec, int64_t neq);
/* vectorize (src/ME_general.f90:1871-1989): device -> host */
int qd_op_download(const qd_op *op, double *vec, int64_t neq);
/* Device address of the operator's neq doubles (for callers that keep their
 * own ODE state on the GPU); valid until qd_op_destroy. */
double *qd_op_device_ptr(qd_op *op);

/* ---- cross-coupled matrices: type(cc_mat), copy_cc ---------------------
 * src/IMSRG_tools.f90:437-461 (copy_cc: zeroed X). */
int qd_cc_create(qd_context *ctx, double herm, qd_cc **out);
int qd_cc_destroy(qd_cc *cc);
int qd_cc_set_herm(qd_cc *cc, double herm);  /* cc_mat%herm (real) */
int qd_cc_upload(qd_cc *cc, const double *x, int64_t cc_size);
int qd_cc_download(const qd_cc *cc, double *x, int64_t cc_size);
int qd_cc_zero(qd_context *ctx, qd_cc *cc);

/* ---- commutator entry points (module commutators) ----------------------
 * [r1, r2] pieces with the reference's accumulation contract: 111 overwrites
 * the one-body part of r3, 121 and 221 add to it; 122 overwrites the
 * two-body part, 222 adds to it.  herm flags are read from r1/r2. */
/* commutator_110(m,n,r1,r2,C), src/commutators.f90:12-33 */
int qd_commutator_110(qd_context *ctx, const qd_op *r1, const qd_op *r2, double *c);
/* commutator_220(r1,r2,c), src/commutators.f90:36-70 */
int qd_commutator_220(qd_context *ctx, const qd_op *r1, const qd_op *r2, double *c);
/* commutator_111(r1,r2,r3), src/commutators.f90:137-195 */
int qd_commutator_111(qd_context *ctx, const qd_op *r1, const qd_op *r2, qd_op *r3);
/* commutator_121(r1,r2,r3), src/commutators.f90:73-134 */
int qd_commutator_121(qd_context *ctx, const qd_op *r1, const qd_op *r2, qd_op *r3);
/* commutator_221(r1,r2,r3,w1,w2), src/commutators.f90:198-352; w1/w2 are
 * context workspace */
int qd_commutator_221(qd_context *ctx, const qd_op *r1, const qd_op *r2, qd_op *r3);
/* commutator_122(r1,r2,r3), src/commutators.f90:355-676.  The two-body pass
 * takes <J|Z|I> as h_Z <I|Z|J> (h_Z = -h1 h2) where one staged class-pair box
 * holds both, which assumes operators that obey their herm flags (square
 * sub-blocks and fhh, fpp (anti)symmetric); equal to the element-by-element
 * evaluation to rounding. */
int qd_commutator_122(qd_context *ctx, const qd_op *r1, const qd_op *r2, qd_op *r3);
/* commutator_222(r1,r2,r3,w1,HCC,ETACC), src/commutators.f90:679-1132 with
 * construct_cc_ints (src/ME_general.f90:923-952): hcc = CC of r2, etacc = CC
 * of r1, both supplied by the caller (zero in the reference's Magnus driver,
 * calc_cc-filled in its ODE RHS). */
int qd_commutator_222(qd_context *ctx, const qd_op *r1, const qd_op *r2, qd_op *r3,
                      const qd_cc *hcc, const qd_cc *etacc);
/* calc_cc(H,HCC), src/commutators.f90:1135-1159 */
int qd_calc_cc(qd_context *ctx, const qd_op *h, qd_cc *hcc);
/* build_white(gen,rec), src/IMSRG_tools.f90:377-434: writes gen's fph and
 * Vpphh only (the caller sets gen herm = -1). */
int qd_build_white(qd_context *ctx, qd_op *gen, const qd_op *rec);

/* ---- operator algebra (src/ME_general.f90:592-686, 2112-2138) ---------- */
/* add_arrays(r1,s1,r2,s2,r3): r3 = s1*r1 + s2*r2 (r3 may alias r1 or r2) */
int qd_add_arrays(qd_context *ctx, const qd_op *r1, double s1, const qd_op *r2, double s2, qd_op *r3);
/* copy_arrays(r1,r2): r2 = r1 (values only; herm is not copied, as in the reference) */
int qd_copy_arrays(qd_context *ctx, const qd_op *r1, qd_op *r2);
int qd_set_to_zero(qd_context *ctx, qd_op *r);
/* mat_2_norm(r): Frobenius norm of every entry except E0 */
int qd_mat_2_norm(qd_context *ctx, const qd_op *r, double *out);
/* MBPT2(rec), src/HFQDmod.f90:325-362 */
int qd_mbpt2(qd_context *ctx, const qd_op *h, double *out);
/* qd_mbpt2 for callers that cannot allocate (code under graph capture, tight
 * loops, harnesses that forbid hipMalloc); no counterpart in the reference.
 * Same kernels and result as qd_mbpt2, but the reduction partials and the
 * result live in a file-scope device array of 4096 + 8 doubles (32 KB per
 * GPU, held until the process ends), so the call allocates nothing.  A
 * channel structure needing more than 4096 partials (one per 256 pphh
 * elements; N=42 R=20 needs 2710) is rejected with QD_E_ARG.  All calls on
 * one GPU share that array: calls that overlap on one GPU race. */
int qd_mbpt2_noalloc(qd_context *ctx, const qd_op *h, double *out);

/* ---- whole commutator and flow-level entry points ----------------------
 * Full [r1,r2] -> r3 (0b in r3's E0, also returned in *e0 if non-null), in
 * the call order of BCH_EXPAND (src/magnus_IMSRG.f90:297-311) and dGam
 * (src/gs_IMSRG.f90:256-268).  hcc/etacc as for qd_commutator_222; NULL
 * means an all-zero CC (the reference's Magnus driver). */
int qd_commutator(qd_context *ctx, const qd_op *r1, const qd_op *r2, qd_op *r3,
                  const qd_cc *hcc, const qd_cc *etacc, double *e0);
/* calc_cc(r2, hcc) and calc_cc(r1, etacc), then the full [r1,r2] -> r3 with
 * them: the sequence of dGam (src/gs_IMSRG.f90:256-268) in one call.  Both CC
 * sets are overwritten (their herm fields are kept).  In the FP64 grouped form
 * the two calc_cc launches overlap the ladder products, which do not read them.
 * Equal to qd_calc_cc twice and qd_commutator to rounding (1e-14): when the CC
 * herm fields equal the operators' flags, the two-body pass takes <J|Z|I> as
 * h_Z <I|Z|J> (h_Z = -h1 h2) for its 1b x 2b and Pandya parts, which assumes
 * operators that obey their herm flags (square sub-blocks and fhh, fpp
 * (anti)symmetric), as every full_ham of the reference does.  qd_dgam,
 * qd_bch_expand and every call without the Pandya term do the same;
 * qd_commutator and qd_commutator_222 with caller-supplied CC sets evaluate
 * each stored element on its own. */
int qd_commutator_cc(qd_context *ctx, const qd_op *r1, const qd_op *r2, qd_op *r3,
                     qd_cc *hcc, qd_cc *etacc, double *e0);
/* BCH_EXPAND(HS,G,H,...), src/magnus_IMSRG.f90:253-318: hs = sum_k cof(k)
 * ad_G^k H with the reference's literal coefficients and its 1e-5 (REAL(4))
 * relative truncation; *nterms = nested commutators taken (1..11).  G is
 * scanned once for sub-block types that are zero in every block (a Magnus
 * Omega holds only fph and Vpphh); the GEMM products reading them are exact
 * zeros and are not launched (QD_ZSKIP=0 launches them). */
int qd_bch_expand(qd_context *ctx, qd_op *hs, const qd_op *g, const qd_op *h,
                  const qd_cc *hcc, const qd_cc *occ, int *nterms);
/* MAGNUS_EXPAND(DG,G,ETA,...), src/magnus_IMSRG.f90:321-369, as shipped:
 * dOmega = eta (the routine returns at :345 before its nested sum). */
int qd_magnus_expand(qd_context *ctx, qd_op *dg, const qd_op *g, const qd_op *eta);
/* euler_step(G,DG,s,stp), src/magnus_IMSRG.f90:372-384: G += stp*DG, s += stp */
int qd_euler_step(qd_context *ctx, qd_op *g, const qd_op *dg, double *s, double stp);
/* dGam(t,yp,H,HCC,ETACC), src/gs_IMSRG.f90:223-273: der = [eta_White(H), H]
 * with calc_cc of H into hcc and of eta into etacc (both overwritten). */
int qd_dgam(qd_context *ctx, const qd_op *h, qd_op *der, qd_cc *hcc, qd_cc *etacc);

/* s_evolve, src/magnus_IMSRG.f90:194-248: the Magnus flow with the White
 * generator from H(0) = h0, Euler steps on Omega, BCH for H(s), reject-and-
 * halve when ||eta|| grows past 1.5x its initial value, stop when
 * |E_MBPT2(H(s))| <= conv (or, with stop_rule 1, |dE/ds| <= conv_deds).  Literal constants are the reference's REAL(4)
 * values widened to REAL(8).  pandya = 0 reproduces the reference as shipped
 * (zero CC inputs); 1 fills the CC matrices with calc_cc every commutator. */
typedef struct {
  double stp;        /* initial step (reference: REAL(4) 0.7)                      */
  double conv;       /* stop when |MBPT2(H(s))| <= conv (reference: REAL(4) 1e-9)   */
  int max_iter;      /* safety cap on accepted + rejected steps                     */
  int pandya;        /* 0 as shipped, 1 with calc_cc-filled CC inputs                */
  int verbose;       /* 1: print the reference's per-iteration line to stdout;
                       2: the same line to stderr, flushed every iteration      */
  int warm_start;    /* 1: start from the Omega passed in (H(s) = BCH(Omega, H(0)));
                       the step-rejection scale ||eta_0|| is then taken from
                       H(0), as in a cold start                                 */
  int diis;          /* 0: the reference's iteration.  m > 0 (at most 6): once
                       |MBPT2(H(s))| <= 1e-4, Pulay (DIIS) acceleration of the
                       same map Omega -> Omega + stp eta(Omega) over the last m
                       iterates (history dropped when ||eta|| grows past twice
                       its smallest value since); before that, and after a
                       drop, the reference's step and rejection rule.  Same
                       stopping rule.  No counterpart in the reference; it
                       removes the slow tail of the plain map near decoupling,
                       which is what a warm start lands in.                    */
  int stop_rule;     /* 0: stop when |MBPT2(H(s))| <= conv (the reference).  1: stop
                       when |dE/ds| <= conv_deds, with dE/ds the zero-body part of
                       [eta(s), H(s)] (commutator_110 + commutator_220), the
                       energy derivative of the IM-SRG flow equation.  No
                       counterpart in the reference.                           */
  double conv_deds;  /* threshold of stop_rule 1 (default 1e-8 Ha)                  */
  int tda_ml, tda_ms, tda_cut;  /* TDA block of the per-step spectrum (qd_tda_spectrum;
                                   tda_cut 0: an empty block)                      */
  const char *spectrum_path;    /* non-null: the reference's spectrum_magnus.dat
                       (src/magnus_IMSRG.f90:130-136, 165-189, 232-243): the file is
                       rewritten from the start; one line for H(0), then two per
                       accepted step (write_spec is called twice there), each
                       n(M-n)+2 fields '(f12.7)': s, E0, then E0 + the TDA
                       eigenvalues, the rest 0 (left undefined by the reference) */
} qd_flow_params;
typedef struct {
  double s, e0, crit;          /* final flow parameter, E0(s), |MBPT2(H(s))|          */
  int iterations, rejected;    /* accepted and rejected Euler steps                  */
  int64_t bch_terms;           /* nested commutators evaluated in all BCH expansions  */
  int converged;
  int trace_rows;              /* rows written to trace                              */
  double mbpt2, deds;          /* MBPT2(H(s)) and dE/ds = [eta(s), H(s)]_0 at the end */
} qd_flow_result;
/* Fills *p with the reference's constants. */
void qd_flow_params_default(qd_flow_params *p);
/* Runs the flow.  omega on entry is the starting Omega (zero for a cold
 * start, a predicted Omega for a warm start); on exit the final Omega.
 * hs receives H(s).  trace (optional, capacity trace_len rows of 6 doubles:
 * iteration, s, E0, nrm1, nrm2, crit) receives the per-step log. */
int qd_magnus_flow(qd_context *ctx, const qd_op *h0, qd_op *hs, qd_op *omega,
                   const qd_flow_params *p, qd_flow_result *res, double *trace, int trace_len);

/* ---- the ODE path --------------------------------------------------------
 * The reference's gs_IMSRG integrates dH/ds = dGam(H) (src/gs_IMSRG.f90:
 * 223-273) with the Shampine-Gordon Adams PECE code (src/adams_ode.f90).
 * qd_ode holds that integrator's state (the reference's work/iwork) on the
 * device; qd_ode_integrate is ode(dGam, neq, y, H, HCC, ETACC, t, tout,
 * relerr, abserr, iflag, work, iwork) (src/adams_ode.f90:7) with y a
 * device operator: on iflag = 1 y is the initial value, on return it holds
 * the solution at *t (= tout when iflag = 2).  iflag, relerr, abserr and t
 * follow the reference: 2 success, 3 tolerances raised, 4 over 500 steps,
 * 5 also looks stiff, 6 bad input. */
typedef struct qd_ode qd_ode;
int qd_ode_create(qd_context *ctx, qd_ode **out);
int qd_ode_destroy(qd_ode *o);
int qd_ode_integrate(qd_ode *o, qd_op *y, double *t, double tout, double *relerr, double *abserr, int *iflag);
/* successful steps, right-hand side evaluations, last order, next step size */
int qd_ode_stats(const qd_ode *o, int *nsteps, int64_t *nfe, int *order, double *h);
/* gs_IMSRG's decoupling loop (src/gs_IMSRG.f90:100-120): ode over output
 * intervals of stp until |(E_old - E_new)/E_old| over one interval is below
 * conv.  Defaults are the reference's REAL(4) literals widened: relerr 1e-8,
 * abserr 1e-6, stp 0.1, conv 1e-7.  As in the reference, a call of ode that
 * returns iflag 3 (tolerances raised), 4 (500 steps) or 5 (stiff) is followed
 * by the next call with that flag, which continues the integration; iflag 6
 * (bad input) ends the loop unconverged.  hs receives H(s); trace (optional,
 * trace_len rows of 4: output, s, E0, crit) the reference's per-output line. */
typedef struct {
  double relerr, abserr, stp, conv;
  int max_outputs;
  int verbose;       /* 1: the reference's per-output line to stdout, 2: stderr */
  int maxnum;        /* integrator steps per ode call before iflag 4 (reference: 500) */
} qd_gs_params;
typedef struct {
  double s, e0, crit;
  int outputs, nsteps;
  int64_t nfe;
  int iflag, converged;
} qd_gs_result;
void qd_gs_params_default(qd_gs_params *p);
int qd_gs_flow(qd_context *ctx, const qd_op *h0, qd_op *hs, const qd_gs_params *p, qd_gs_result *res,
               double *trace, int trace_len);

/* ---- batched form: launch strategy of the grouped GEMM ------------------
 * The GEMM problems of all (M_L, M_S) blocks run as one grouped persistent
 * launch (mode 0, default) or as one launch per block problem, the way the
 * reference calls dgemm per block (mode 1).  No counterpart in the reference:
 * this is the batching control of the MI355X engine. */
int qd_set_launch_mode(qd_context *ctx, int mode);
/* 1: qd_bch_expand captures each nested-commutator term (calc_cc of the
 * term, the fused commutator, both norms and the update of hs) in a hipGraph
 * on first use and replays it, and qd_commutator_cc does the same for the
 * whole call; graphs are keyed by the operand addresses, hermiticities, the
 * term index and the zero sub-block mask of G.  0 (default): plain stream
 * launches.  No counterpart in the reference. */
int qd_set_graphs(qd_context *ctx, int on);
/* Number of captured graphs the context holds.  A graph is dropped when an
 * operator or CC set it was captured on is destroyed. */
int64_t qd_graph_count(const qd_context *ctx);
/* Project autotuner: coordinate descent over the (M, N, K) size buckets of
 * the grouped GEMM problems, largest work first.  For each bucket every tile
 * shape (64x64, 128x64, 32x32, 128x16, 64x16, 16x64, 16x128, 64x16 split-K,
 * 128x128) is tried, timing one commutator's whole GEMM work (both grouped
 * lists; the fastest of reps runs) on scratch operands at the current
 * precision, and the fastest is kept.  64x64 and 32x32 tiles share one
 * launch; every other shape present adds a launch.  Reports that device time
 * before and after (ms).  Drops captured graphs. */
int qd_autotune(qd_context *ctx, int reps, double *ms_before, double *ms_after);

/* ---- learned warm start ----------------------------------------------
 * No counterpart in the reference (s_evolve, src/magnus_IMSRG.f90:194-248,
 * always starts from Omega = 0).  A small FP32 multilayer perceptron
 * (nin -> nlayers x nhidden tanh -> nout linear, Glorot-uniform weights from
 * seed, zero biases), trained and evaluated on the context's stream by the
 * library's own kernels, maps frequency features to the coefficients of a PCA
 * basis of converged Omegas; qd_basis_expand turns coefficients into an
 * operator, the warm start of qd_magnus_flow (warm_start = 1).
 * Parameters: one flat FP32 vector, per layer W (out x in, row-major) then b.
 * x: batch x nin, t and y: batch x nout, row-major.  Loss: mean squared
 * error over batch and outputs.  qd_mlp_train: full-batch Adam (beta1 0.9,
 * beta2 0.999, eps 1e-8), optional per-step loss trace (length steps). */
typedef struct qd_mlp qd_mlp;
typedef struct qd_basis qd_basis;
int qd_mlp_create(qd_context *ctx, int nin, int nhidden, int nlayers, int nout, uint64_t seed, qd_mlp **out);
int qd_mlp_destroy(qd_mlp *m);
int64_t qd_mlp_num_params(const qd_mlp *m);
int qd_mlp_get_params(qd_mlp *m, float *p, int64_t n);
int qd_mlp_set_params(qd_mlp *m, const float *p, int64_t n);  /* also resets the Adam state */
int qd_mlp_forward(qd_mlp *m, const float *x, int batch, float *y);
int qd_mlp_loss_grad(qd_mlp *m, const float *x, const float *t, int batch, float *loss, float *grad);
int qd_mlp_train(qd_mlp *m, const float *x, const float *t, int batch, int steps, float lr, float *loss_trace);
/* PCA basis of operators: mean (neq) and rank vectors (rank x neq, row-major),
 * FP64, held on the device.  qd_basis_expand: out = mean + sum_r coef[r] vecs[r]. */
int qd_basis_create(qd_context *ctx, int rank, const double *mean, const double *vecs, qd_basis **out);
int qd_basis_destroy(qd_basis *b);
int qd_basis_expand(qd_basis *b, const double *coef, qd_op *out);

/* ---- precision form: arithmetic of the 2b-2b GEMMs ----------------------
 * The GEMMs of commutator_222 (matrix_mult222 products and accumulations,
 * the Pandya products of construct_cc_ints) and the commutator_221
 * intermediates they share:
 *   QD_PREC_FP64    FP64 MFMA (default; the plain form)
 *   QD_PREC_FP32    operands rounded to FP32, FP32 MFMA
 *   QD_PREC_BF16X2..4  operands cut into 2..4 BF16 mantissa slices; slice
 *                   pairs (i, j) with i + j < S on BF16 MFMA
 * Split modes accumulate in FP32 over K stages of 32 and in FP64 across
 * stages and into the output.  Everything else stays FP64: the one-body
 * pieces, 122, the gather of the Pandya term, the zero-body traces, White denominators,
 * the BCH sum, norms and the Omega update.  No counterpart in the reference
 * (src/commutators.f90:999-1127 and src/ME_general.f90:948-949 call dgemm).
 * Drops captured BCH graphs. */
#define QD_PREC_FP64 0
#define QD_PREC_FP32 1
#define QD_PREC_BF16X2 2
#define QD_PREC_BF16X3 3
#define QD_PREC_BF16X4 4
int qd_set_precision(qd_context *ctx, int prec);
int qd_get_precision(qd_context *ctx, int *prec);
/* Current tile shape of every GEMM problem (products, then accumulating);
 * returns the count. */
int qd_gemm_shapes(const qd_context *ctx, int *shapes, int64_t len);
/* Sets every problem's tile shape (len = the count qd_gemm_shapes returns,
 * values 0..8 as it reports them), e.g. to replay an autotuned set in another
 * process; drops captured BCH graphs.  No counterpart in the reference. */
int qd_set_gemm_shapes(qd_context *ctx, const int *shapes, int64_t len);

/* ---- input side (host FP64, no GPU) --------------------------------------
 * v_int(qi,qj,qk,ql,hw), src/ME_general.f90:2205-2276: the analytic 2D
 * oscillator Coulomb element <i j|V|l k> for (n, m_l, m_s) triples. */
double qd_v_int(const int *qi, const int *qj, const int *qk, const int *ql, double hw);
/* construct_two_particle_HF_BASIS + HF_E2 (src/HFQDmod.f90:9-254, 292-323):
 * H(0) in the Hartree-Fock basis, written to h0 (neq doubles, vectorize
 * order; E0 = E_HF).  The interaction is built from v_int as get_ints
 * (src/ME_general.f90:746-798) does instead of being read from ME.bin.  eig
 * (optional, M doubles) receives the HF single-particle energies in states
 * order, iters the SCF iteration count.  Eigenvector phases: the
 * largest-magnitude component of each HF orbital is positive. */
int qd_hf_hamiltonian(int nbody, int nshell, double hw, double *h0, int64_t neq, double *eig, int *iters);
/* The same H(0) with the interaction read from an ME.bin file, as
 * construct_two_particle_HF_BASIS reads ../TBMEfiles/ME.bin (src/HFQDmod.f90:
 * 97-129): a stream of records n1,m1,n2,m2,n3,m3,n4,m4 (int32) and Vdir
 * (real64) = <ab|V|cd> at hw = 1, scaled by sqrt(hw) on load; records with an
 * orbital beyond nshell shells are skipped; each record sets <ab|V|cd> and
 * <cd|V|ab>; the exchange element falls back to <ba|V|cd> where <ab|V|dc> is
 * below 1e-10 (src/HFQDmod.f90:163-166).  A truncated record is an error (the
 * reference's 'failfile' stop). */
int qd_hf_hamiltonian_me(int nbody, int nshell, double hw, const char *me_path, double *h0, int64_t neq,
                         double *eig, int *iters);
/* qd_hf_hamiltonian with the heavy parts on the device and H(0) written into
 * the hermitian operator h0 of ctx (its layout gives nbody and nshell): the
 * Fock build of every SCF iteration (construct_V, src/HFQDmod.f90:196-221,
 * 260-290) and the two-body basis transform (new_VMAT, src/ME_general.f90:
 * 1362-1492) run on the context's stream from the Coulomb table uploaded once;
 * the per-channel diagonalisations and the convergence test stay on the host.
 * Same SCF, phases and E0 rule as qd_hf_hamiltonian; results agree with it to
 * FP64 rounding. */
int qd_hf_hamiltonian_device(qd_context *ctx, double hw, qd_op *h0, double *eig, int *iters);
/* Writes ME.bin for nshell shells from v_int at hw = 1: one record per
 * (a, b, c, d) with m_a + m_b = m_c + m_d and (a, b) <= (c, d) in orbital
 * order; *nrec (optional) receives the record count.  The reference ships no
 * writer (its ME.bin is an external file). */
int qd_me_bin_write(int nshell, const char *me_path, int64_t *nrec);

/* ---- consumers of the decoupled H(s) --------------------------------------
 * TDA spectrum: build_TDAmat + calc_TDA + diagonalize_blocks (src/operators.f90:
 * 7-124, src/ME_general.f90:576-590) for the (M_L, M_S) = (ml, ms) transfer
 * (m_s in units of 1/2, as the reference's states) with particles below the
 * 1-based state cut (a = N+1 .. cut): the matrix d_ij f_ab - d_ab f_ji +
 * <aj|V|ib> over (hole i, particle a) pairs, hole outer, gathered on the
 * device; its eigenvalues (ascending, relative to E0) on the host from the
 * upper triangle, as dsyev('V','U') reads it.  *dim receives the block
 * dimension (0 for an empty block); eig needs that many entries. */
int qd_tda_spectrum(qd_context *ctx, const qd_op *h, int ml, int ms, int cut, double *eig, int cap, int *dim);
/* EOM-IM-SRG excited states of H(s) for the (ml, ms) transfer:
 * calculate_excited_states / LANCZOS_DIAGONALIZE / matvec_prod
 * (src/EOM_IMSRG.f90:10-47, 132-280, 551-660) with HQ_comm_1b / HQ_comm_2b
 * (src/commutators.f90:1162-1260).  The ladder operator has ph components
 * (particle outer) then pphh components (a < b, i < j, a outer), the order of
 * unwrap/rewrap (src/EOM_IMSRG.f90:815-884, 1011-1080).
 * qd_eom_dim: the counts of ph and pphh components.
 * qd_eom_matvec: y = [H, Q(x)] on those components (host vectors of length n).
 * qd_eom_excited_states: the nev smallest eigenvalues (excitation energies
 * dE, ascending) by Lanczos with full reorthogonalisation on the device (in
 * place of ARPACK's dsaupd 'SA'; relative residual 1e-10), the ph weight
 * sum(fph**2) of each eigenvector (optional) and the Lanczos steps taken. */
int qd_eom_dim(const qd_context *ctx, int ml, int ms, int64_t *n_ph, int64_t *n_pphh);
int qd_eom_matvec(qd_context *ctx, const qd_op *h, int ml, int ms, const double *x, double *y, int64_t n);
int qd_eom_excited_states(qd_context *ctx, const qd_op *h, int ml, int ms, int nev, double *eig, double *w1p1h,
                          int *iters);
/* One particle attached (mode 1) or removed (mode 2): calculate_1p_attached /
 * calculate_1h_removed with LANCZOS_1p_ATTACHED / LANCZOS_1h_REMOVED and
 * matvec_1p_attached / matvec_1h_removed (src/EOM_IMSRG.f90:50-128, 283-548,
 * 662-812; HQ_comm_1p/2p1h/1h/2h1p, src/commutators.f90:1263-1425).  The
 * ladder operator has one-index components (mode 1: particles a with
 * (m_l, m_s) = (ml, ms); mode 2: holes i with -(m_l, m_s) = (ml, ms)), then
 * three-index ones (mode 1: a < b particles, hole i; mode 2: i < j holes,
 * particle a), the order of unwrap_1p / unwrap_1h (src/EOM_IMSRG.f90:886-1009).
 * qd_eom13_dim: the two counts; qd_eom13_matvec: y = A x (host vectors);
 * qd_eom_attached_removed: the nev smallest eigenvalues dE, the weight of the
 * one-index components of each eigenvector (norm_1p, src/EOM_IMSRG.f90:409-
 * 411) and the Lanczos steps. */
int qd_eom13_dim(const qd_context *ctx, int mode, int ml, int ms, int64_t *n1, int64_t *n3);
int qd_eom13_matvec(qd_context *ctx, const qd_op *h, int mode, int ml, int ms, const double *x, double *y, int64_t n);
int qd_eom_attached_removed(qd_context *ctx, const qd_op *h, int mode, int ml, int ms, int nev, double *eig,
                            double *w1, int *iters);

/* ---- on-disk format of an operator --------------------------------------
 * export_hamiltonian / import_hamiltonian (src/ME_general.f90:207-361): the
 * file the reference's gs_IMSRG writes to ../hamiltonians/ (src/gs_IMSRG.f90:
 * 141-142), Fortran unformatted sequential with one record per REAL(8)
 * (4-byte length 8, value, length): E0, fhh, fph, fpp, then per block Vhhhh,
 * Vpppp, Vpphh, Vphph, Vphhh, Vppph, every matrix row by row.  path is the
 * whole file name (the caller adds the reference's ../hamiltonians/).
 * Import checks the size and every record length.  The _host forms take a
 * vectorize-order host vector and need no GPU. */
int qd_export_hamiltonian(const qd_op *op, const char *path);
int qd_import_hamiltonian(qd_op *op, const char *path);
int qd_export_hamiltonian_host(int nbody, int nshell, const double *vec, int64_t neq, const char *path);
int qd_import_hamiltonian_host(int nbody, int nshell, double *vec, int64_t neq, const char *path);

/* ---- timing of the dominant kernels (HIP events on the context stream) --
 * Accumulates per-kernel-class device time while enabled.  Classes: 0 the
 * product GEMM launch (ladder intermediates + Pandya products), 1 the
 * accumulating ladder GEMM launch, 2 two-body pass (122, ladder combine and
 * the gather of the Pandya term), 3 one-body, 4 calc_cc, 5 other, 6 reserved
 * (always 0; the Pandya gather joined class 2). */
int qd_timing_enable(qd_context *ctx, int on);
int qd_timing_read(qd_context *ctx, double *ms_per_class, int64_t *launches_per_class, int nclass);
/* Algorithmic FP64 flops of one full commutator's GEMM launches: the product
 * launch without / with its Pandya part, and the accumulating launch. */
int qd_gemm_flops(const qd_context *ctx, double *ladder, double *pandya, double *acc);
/* Matrix-core flops the same launches execute at the current tile shapes and
 * precision (zero-filled tile edges and K steps included; BF16 modes count
 * every slice product, FP32 and BF16 full K stages of 32): the product
 * launch's ladder and one-body problems, its Pandya part (the fused kernel's
 * tile pairs in the grouped FP64 and BF16 forms), and the accumulating
 * launch.  Against
 * qd_gemm_flops this is the padding of each list; against the launch time,
 * its matrix-core rate. */
int qd_gemm_mfma_flops(const qd_context *ctx, double *products, double *pandya, double *acc);

#ifdef __cplusplus
}
#endif
#endif /* QDIMSRG_H */
