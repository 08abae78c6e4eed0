/*
 * fgt_tree.h -- host adaptive-tree builder of libfgt_amd.so.
 *
 * Replaces the reference's two-pass builder vol_tree_mem / vol_tree_build
 * (src/common/tree_vol_coeffs_nd.f:83-437, 444-687) for a source given as a
 * sum of anisotropic Gaussians, with the same refinement rule (L2 tail of
 * the tensor-product coefficients, iptype = 2, eta = 0), the same 2:1 level
 * restriction (vol_tree_fix_lr, 1242-1557) and the same output arrays
 * (itree/iptr/centers/boxsize/fvals) that boxfgt_ consumes.
 */
#ifndef FGT_AMD_FGT_TREE_H
#define FGT_AMD_FGT_TREE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct fgt_tree_builder fgt_tree_builder;

/* A source given by the caller, with the calling convention of the
 * reference's `fun` (tree_vol_coeffs_nd.f:83-85): every argument by
 * reference; f receives nd values at the point xyz(ndim).  It is called from
 * several threads at once. */
typedef void (*fgt_source_fn)(int *nd, double *xyz, double *dpars, void *zpars, int *ipars, double *f);

/* First pass: builds the tree in an internal buffer and reports its sizes
 * (vol_tree_mem).  gauss: ng rows of (c[ndim], s[ndim], a); the source is
 * f(y) = sum a exp(-sum_k (y_k - c_k)^2 / s_k), periodised over the 3^ndim
 * nearest images when iperiod = 1.  The root box is [-boxlen/2, boxlen/2]^ndim. */
int fgt_vol_tree_mem(int ndim, int ipoly, int iperiod, double eps, int norder,
                     double boxlen, int ng, const double *gauss,
                     fgt_tree_builder **out, int *nboxes, int *nlevels,
                     int *ltree);
/* fgt_vol_tree_mem with a size limit (no counterpart in the reference):
 * returns status 3 without a builder as soon as the tree would have more
 * than max_leaves leaves (0 = no limit). */
int fgt_vol_tree_mem_capped(int ndim, int ipoly, int iperiod, double eps, int norder,
                            double boxlen, int ng, const double *gauss,
                            long long max_leaves, fgt_tree_builder **out,
                            int *nboxes, int *nlevels, int *ltree);
/* vol_tree_mem for any source (tree_vol_coeffs_nd.f:83-437): fun with its
 * dpars / zpars / ipars, nd densities, refinement in the iptype norm
 * (0 max, 1 L1, 2 L2) of the tail coefficients scaled by (box size / 2)^eta
 * against eps times the running norm of the source, boxes with
 * zk * size > 5 always refined; ifnewtree = 1 reports three times the boxes
 * (room for boxfgt's refinement).  Same arrays as the reference builds. */
int fgt_vol_tree_mem_fun(int ndim, int ipoly, int iperiod, double eps, double zk,
                         double boxlen, int norder, int iptype, double eta,
                         fgt_source_fn fun, int nd, double *dpars, void *zpars,
                         int *ipars, int ifnewtree, fgt_tree_builder **out,
                         int *nboxes, int *nlevels, int *ltree);
/* Room for boxfgt's refinement on a built tree (vol_tree_mem with
 * ifnewtree = 1): the arrays fgt_vol_tree_build fills become three times
 * the boxes; new sizes in *nboxes, *ltree. */
int fgt_vol_tree_newtree_room(fgt_tree_builder *b, int *nboxes, int *ltree);
/* Second pass: copies the arrays out (vol_tree_build) and frees the builder.
 * itree(ltree), iptr(8), centers(ndim,nboxes), boxsize(0:nlevels),
 * fvals(nd,npbox,nboxes), nboxes and ltree as the first pass reported them
 * (boxes past the tree's own are zero). */
int fgt_vol_tree_build(fgt_tree_builder *b, int *itree, int *iptr,
                       double *centers, double *boxsize, double *fvals);

#ifdef __cplusplus
}
#endif
#endif
