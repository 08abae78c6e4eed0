/*
 * fgt.h -- C ABI of the MI355X box Gauss transform (libfgt_amd.so).
 *
 * Drop-in boundary: the Fortran external subroutine boxfgt of
 * flatironinstitute/fgt (src/bfgt/boxfgt.f:5-8, documented at README.md:61-63).
 * The tree is built on the host beforehand, by the reference's
 * vol_tree_mem/vol_tree_build (src/common/tree_vol_coeffs_nd.f:83-85,
 * 444-446) or by fgt_vol_tree_mem/fgt_vol_tree_build below, which produce
 * the same arrays.
 *
 * All integers are 32-bit (default Fortran INTEGER), all reals FP64, all
 * arrays are in Fortran (column-major) layout, box ids in itree are 1-based.
 */
#ifndef FGT_AMD_FGT_H
#define FGT_AMD_FGT_H
#ifdef __cplusplus
extern "C" {
#endif

/*
 * boxfgt_ -- replaces boxfgt (src/bfgt/boxfgt.f:5-280).  Same 28 by-reference
 * arguments, gfortran/flang default mangling, so a Fortran caller links it
 * in place of the reference.  pot(nd,npbox,nboxes) is accumulated (+=) at
 * the tensor nodes of every leaf, as the reference does; the caller zeroes it.
 * ifpgh = 2 / 3 also accumulate grad(nd,ndim,npbox,nboxes) and
 * hess(nd,ndim*(ndim+1)/2,npbox,nboxes) (Hessian order 1-D xx; 2-D xx,xy,yy;
 * 3-D xx,yy,zz,xy,xz,yz).
 * ifpghtarg = 1..3 with ntarg >= 1 evaluates pot (+ grad, + hess) at the
 * targets targs(ndim,ntarg) into pote(nd,ntarg), grade(nd,ndim,ntarg),
 * hesse(nd,nhess,ntarg) (overwritten), from the polynomial expansions of
 * the leaf values, as the reference's STEP 11 (boxfgt.f:1343-1378).
 * ifnewtree = 1 rewrites the tree after the solve, as the reference's STEPs
 * 8-10 (boxfgt.f:1026-1337): leaves whose potential is under-resolved are
 * refined (the potential, and grad / hess, of each new box evaluated from its
 * parent's local expansion and near-field list), sibling leaves whose merged
 * potential is resolved are coarsened when the refinement added a tenth of
 * the boxes, and the 2:1 level restriction is restored with interpolated
 * values; itree, centers, *nlevels and pot (grad, hess) are rewritten and the
 * targets are evaluated on the new tree.  nboxes is then the capacity of the
 * arrays (vol_tree_mem with ifnewtree = 1 gives three times the tree's
 * boxes; the tree's own boxes are 1..laddr(2,nlevels)); a refinement that
 * needs more returns status 6 with the outputs partly rewritten.  Boxes at
 * the finest level are not refined again (the reference's level arrays end
 * there).  Rows past the new box count are left unspecified.
 * Supported: ifpgh = 1..3, ifpghtarg = 0..3, ifnewtree = 0, 1.  Anything else
 * leaves the outputs untouched and records a nonzero status
 * (fgt_last_status); the reference has no error channel (it neither checks
 * nor reports).
 * timeinfo(1..20) is zeroed; (1) plan+upload, (2..7) STEP 1..6 device
 * seconds, (8) download, (9) STEPs 8-10, (12) extra targets.
 */
void boxfgt_(int *nd, int *ndim, double *delta, double *eps, int *ipoly,
             int *iperiod, int *norder, int *npbox, int *nboxes, int *nlevels,
             int *ltree, int *itree, int *iptr, double *centers,
             double *boxsize, double *fvals, int *ifpgh, double *pot,
             double *grad, double *hess, int *ifnewtree, int *ntarg,
             double *targs, int *ifpghtarg, double *pote, double *grade,
             double *hesse, double *timeinfo);

/* fgt_boxfgt -- the same entry point under a bind(C) name, for the
 * iso_c_binding interface block in INTEGRATION.md (replaces the same
 * src/bfgt/boxfgt.f:5-8).  Returns the status as well. */
int fgt_boxfgt(int *nd, int *ndim, double *delta, double *eps, int *ipoly,
               int *iperiod, int *norder, int *npbox, int *nboxes,
               int *nlevels, int *ltree, int *itree, int *iptr,
               double *centers, double *boxsize, double *fvals, int *ifpgh,
               double *pot, double *grad, double *hess, int *ifnewtree,
               int *ntarg, double *targs, int *ifpghtarg, double *pote,
               double *grade, double *hesse, double *timeinfo);

/*
 * fgt_boxfgt_noalloc -- boxfgt with device-resident data that allocates no
 * device memory.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same computation as boxfgt_ with
 * ifpgh = 1, ifpghtarg = 0, ifnewtree = 0: fvals(nd,npbox,nboxes) and
 * pot(nd,npbox,nboxes) are DEVICE pointers and pot is accumulated (+=) on
 * `stream` (a hipStream_t; NULL = default stream); the tree arrays are host
 * arrays, planned on the host.  Every other device buffer (tables, work
 * lists, plane-wave expansions) is carved from a 128 MiB per-GPU static
 * device workspace; a problem that needs more returns status 2.  Calls are
 * synchronous (they return when the result is in pot).  All calls on one GPU
 * share that workspace: calls that overlap in time on one GPU race.
 */
int fgt_boxfgt_noalloc(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder,
                       int nboxes, int nlevels, int ltree, const int *itree, const int *iptr,
                       const double *centers, const double *boxsize, const double *fvals, double *pot,
                       void *stream);

/* boxfgt_ / fgt_boxfgt keep the plan and device buffers of their last few
 * trees (up to 4 entries, 32 GiB): a repeated call on the same tree arrays
 * and parameters skips planning and allocation and moves only fvals and the
 * outputs.  Entries are matched by comparing the stored arrays, not only a
 * hash.  Counters (nullable outputs) and a way to free the device memory
 * they hold.  No counterpart in the reference, which replans every call. */
int fgt_dropin_cache_stats(long long *hits, long long *misses, int *entries);
int fgt_dropin_cache_clear(void);

/* Status of the last call made on this thread: 0 = ok.  Copies the message
 * (NUL-terminated, truncated to len) into msg when msg != NULL.  Added error
 * channel; the reference has none (boxfgt.f:191 ignores its allocation
 * status). */
int fgt_last_status(char *msg, int len);

/*
 * Device-resident solver: the same computation as boxfgt_ split into plan
 * (host lists and tables, device upload), data upload, execution and
 * download, so that repeated solves on one tree keep everything in HBM.
 * It has no single counterpart in the reference: it is bfgtmain
 * (src/bfgt/boxfgt.f:285-1435) with its STEP 0 hoisted out of the solve.
 */
typedef struct fgt_solver fgt_solver;

typedef struct {
    int npwlevel, npw, nexp, nlevstart, nslots;
    int periodic_small, all_pw;
    long long nleaves, npw_leaves, ndirect_pairs, nout; /* nout = nd*npbox*nleaves */
    long long expansion_bytes;
    double stage_ms[8];   /* last execute: STEP 1..6 in [0..5] (when timed) */
    long long n_m2l_targets, n_m2l_pairs;   /* cutoff-level boxes receiving a loc; list entries */
    long long n_parent_child;                /* child->parent merge edges */
    long long m2l_bricks;                    /* STEP 3 sibling bricks, -1 when it runs on target groups */
    int m2l_form;                            /* STEP 3 kernel: 0 target groups, 1 brick entry streams,
                                                2 separable brick stencil */
} fgt_solver_stats;

int fgt_solver_create(int nd, int ndim, double delta, double eps, int ipoly,
                      int iperiod, int norder, int nboxes, int nlevels,
                      int ltree, const int *itree, const int *iptr,
                      const double *centers, const double *boxsize,
                      fgt_solver **out);
/* fvals(nd,npbox,nboxes) host -> device (after all queued device work). */
int fgt_solver_set_fvals(fgt_solver *s, const double *fvals);
/* Runs STEPs 1-6 on `stream` (a hipStream_t; NULL = default stream) into the
 * device potential (overwritten, not accumulated).  timed != 0 records
 * per-stage events (synchronises at the end). */
int fgt_solver_execute(fgt_solver *s, void *stream, int timed);
/* device potential -> host; accumulate != 0 adds into pot (reference +=).
 * Waits for all queued device work first (an execute on any stream). */
int fgt_solver_get_pot(fgt_solver *s, double *pot, int accumulate);
/* Output set of later executes, as boxfgt's ifpgh (src/bfgt/boxfgt.f:35-38):
 * 1 potential, 2 + gradient, 3 + Hessian.  Allocates the device gradient
 * grad(nd,ndim,npbox,nboxes) / Hessian hess(nd,ndim*(ndim+1)/2,npbox,nboxes)
 * arrays (Hessian order 1-D xx; 2-D xx,xy,yy; 3-D xx,yy,zz,xy,xz,yz, as
 * src/bfgt/bfgt_pwrouts.f:612-1065).  Default 1. */
int fgt_solver_set_ifpgh(fgt_solver *s, int ifpgh);
/* Precision of later executes.  No counterpart in the reference (FP64 only).
 *   0  FP64 (default).
 *   1  FP32: plane-wave expansions stored as FP32 complex and the large
 *      tensor-product contractions (STEP 5 x stage) on the FP32 matrix
 *      cores with FP32 accumulation; table and phase values are formed in
 *      FP64, translations (STEPs 2-4) multiply in FP64 and the output
 *      accumulation is FP64.  Needs eps >= 1e-6.
 *   2  BF16 hi/lo: as 1, with those contractions as three BF16 matrix-core
 *      products per FP32 pair (hi*hi + hi*lo + lo*hi).  Needs eps >= 1e-3.
 * Modes 1 and 2 need the 3-D register-blocked path (ndim 3, norder
 * 8/12/16) and a plan that forms plane-wave expansions.  A mode the
 * solver's eps or shape does not allow returns status 2 and changes
 * nothing. */
int fgt_solver_set_precision(fgt_solver *s, int mode);
/* device gradient / Hessian -> host (accumulate != 0 adds, reference +=);
 * status 2 when not requested by fgt_solver_set_ifpgh. */
int fgt_solver_get_grad(fgt_solver *s, double *grad, int accumulate);
int fgt_solver_get_hess(fgt_solver *s, double *hess, int accumulate);
/* STEP 11 on the last execute's device pot / grad / hess: values at ntarg
 * targets targs(ndim,ntarg) (treedata_trans_nd + treedata_evalt_nd,
 * src/common/tree_data_routs_nd.f:16-95,582-686) into pote(nd,ntarg),
 * grade(nd,ndim,ntarg) (ifpghtarg >= 2), hesse(nd,nhess,ntarg) (= 3),
 * overwritten.  Needs ifpghtarg <= the solver's ifpgh. */
int fgt_solver_eval_targets(fgt_solver *s, int ntarg, const double *targs,
                            int ifpghtarg, double *pote, double *grade,
                            double *hesse);
int fgt_solver_stats_get(const fgt_solver *s, fgt_solver_stats *out);
/* Device pointers of the resident fvals / pot buffers (for callers that
 * fill or read HBM directly). */
void *fgt_solver_fvals_device(fgt_solver *s);
void *fgt_solver_pot_device(fgt_solver *s);
void *fgt_solver_grad_device(fgt_solver *s);   /* NULL until ifpgh >= 2 */
void *fgt_solver_hess_device(fgt_solver *s);   /* NULL until ifpgh >= 3 */
void fgt_solver_destroy(fgt_solver *s);

/* ---- many independent problems (no counterpart in the reference, which
 *      solves one problem per call) ---- */
typedef struct {
    int nd, ndim;
    double delta, eps;
    int ipoly, iperiod, norder, nboxes, nlevels, ltree;
    const int *itree, *iptr;
    const double *centers, *boxsize;
} fgt_problem;
typedef struct fgt_batch fgt_batch;
/* Plans every problem (arguments as fgt_solver_create).  3-D problems of
 * norder 8, 12 or 16 are grouped into shape classes (norder, nd, precision,
 * plane-wave count rounded up to 16) and every stage of a class runs as one
 * launch over the tasks of all its problems (STEPs 2 and 4 in level phases
 * aligned per problem); the other problems run their own executes on
 * nstreams (1..16) side streams.  Replaces the reference's one problem per
 * call with its level-synchronous box loops (src/bfgt/boxfgt.f:711-928). */
int fgt_batch_create(int nprob, const fgt_problem *probs, int nstreams,
                     fgt_batch **out);
/* Per-problem precision (fgt_solver_set_precision modes): mode 0 FP64 for
 * all; 1 FP32 where eps >= 1e-6; 2 BF16 hi/lo where eps >= 1e-3; 3 the
 * cheapest mode each problem's eps allows (BF16 hi/lo for eps >= 1e-3, FP32
 * for eps >= 1e-6).  Problems with eps < eps_min, or off the 3-D
 * register-blocked path, stay FP64.  *nmixed (nullable) = problems not in
 * FP64. */
int fgt_batch_set_precision(fgt_batch *b, int mode, double eps_min,
                            int *nmixed);
int fgt_batch_set_fvals(fgt_batch *b, int i, const double *fvals);
/* Runs STEPs 1-6 of every problem and waits for all.  The first execute
 * (after create, a precision change, set_graph or autotune) captures the
 * whole launch sequence into a HIP graph; later executes replay it. */
int fgt_batch_execute(fgt_batch *b);
/* on != 0: replay from a captured graph (default); 0: enqueue eagerly. */
int fgt_batch_set_graph(fgt_batch *b, int on);
/* Autotuner: per shape class, times each candidate STEP 3 targets-per-block
 * (2, 3, 4, 6, 8) and STEP 1 / STEP 5 x-stage variant (VALU, FP64 or
 * reduced-precision matrix cores) reps times on the batch's own data and
 * keeps the fastest; the choices are remembered for the process per class
 * key and applied to later batches. */
int fgt_batch_autotune(fgt_batch *b, int reps);
/* number of shape classes; info[12] of class k: norder, nd, precision,
 * plane-wave cap, members, STEP 3 group, STEP 1 / STEP 5 variant (-1 =
 * default policy), STEP 5 tasks, near-field targets, STEP 3 elements per
 * thread, STEP 3 over sibling bricks (1) or target groups (0) */
int fgt_batch_nclasses(const fgt_batch *b);
int fgt_batch_class_info(const fgt_batch *b, int k, int *info);
/* batch problem indices of class k (info[4] of them) */
int fgt_batch_class_members(const fgt_batch *b, int k, int *members);
/* One eager execute with HIP events around every stage of every class (the
 * reference's per-step timeinfo, boxfgt.f:1419-1432), all on one stream so
 * each stage is timed alone; then fgt_batch_class_stage_ms gives class k's
 * STEP 1..6 device milliseconds. */
int fgt_batch_profile(fgt_batch *b);
int fgt_batch_class_stage_ms(const fgt_batch *b, int k, double *ms);
int fgt_batch_get_pot(fgt_batch *b, int i, double *pot, int accumulate);
/* total output nodes of the batch; device time of the last execute (ms) */
int fgt_batch_stats(const fgt_batch *b, long long *nout, double *last_ms);
void fgt_batch_destroy(fgt_batch *b);

/* ---- serving (no counterpart in the reference) ----
 * A server keeps problems resident on one GPU under a device-memory budget
 * and runs requests as they arrive on its own HIP streams (up to nstreams
 * at once, overlapping on the device).  A request names its problem by a
 * caller key (equal keys = the same tree and parameters; the arrays of a
 * key's problem must stay valid while the server lives) and brings its
 * source values (host fvals, valid until the request finishes; uploaded
 * when the request starts unless the same pointer and version are already
 * resident).  A problem that is not resident is planned on the host when
 * its request comes up and built when its estimated footprint fits the
 * budget after evicting idle problems least recently used first; otherwise
 * the request waits for running ones to finish, and later requests wait
 * behind it (a request whose problem is running another request may be
 * overtaken).  Result buffers count against the budget too.  precision_mode: as fgt_batch_set_precision, per problem.  Each
 * finished request's potential (nd,npbox,nboxes) stays on the device until
 * fgt_server_result copies it out or fgt_server_release drops it. */
typedef struct fgt_server fgt_server;
typedef struct {
    double budget_bytes, used_bytes;
    long long problems, resident, requests, completed, waiting, running;
    long long plans, evictions, mem_waits;
} fgt_server_info;
int fgt_server_create(double budget_bytes, int nstreams, int precision_mode, fgt_server **out);
int fgt_server_submit(fgt_server *v, const fgt_problem *p, long long key, const double *fvals,
                      long long fvals_version, int *req_id);
/* advances the server (collects finished requests, admits and starts
 * waiting ones); done[] receives up to max_done ids finished since the last
 * poll */
int fgt_server_poll(fgt_server *v, int *done, int max_done, int *ndone);
/* polls until no request is waiting or running (status 6 after timeout_s) */
int fgt_server_drain(fgt_server *v, double timeout_s);
/* times[3]: submit, start, finish (seconds on the server's clock, -1 if
 * not yet); flags: bit0 plan-cache hit, bit1 planned for this request, bit2
 * waited for memory, bit3 values uploaded, bits 8-9 state (0 waiting, 1
 * running, 2 done, 3 failed) */
int fgt_server_request_info(const fgt_server *v, int id, double *times, int *flags);
int fgt_server_result(fgt_server *v, int id, double *pot);
int fgt_server_release(fgt_server *v, int id);
int fgt_server_stats(const fgt_server *v, fgt_server_info *out);
void fgt_server_destroy(fgt_server *v);

#ifdef __cplusplus
}
#endif
#endif
