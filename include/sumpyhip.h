/*
 * C-ABI of libsumpyhip.so: hand-written CDNA4 (gfx950) kernels for sumpy's
 * FMM/QBX interaction path.
 *
 * Conventions (all entry points):
 *   - every pointer is a device pointer (caller-owned; the library never
 *     frees caller memory), except the sumpyhip_expn descriptor itself, which
 *     is a host struct whose pointer members are device pointers;
 *   - coordinates are structure-of-arrays: one pointer per axis;
 *   - real kernels use double arrays; the Helmholtz kernel uses interleaved
 *     complex128 (re, im) for weights and outputs;
 *   - "stream" is a hipStream_t; work is enqueued asynchronously on it;
 *   - return value 0 = success, otherwise an error code;
 *     sumpyhip_last_error() describes the last failure of the calling thread.
 *
 * Kernel ids: 2 = Laplace 2D, 3 = Laplace 3D, 13 = Helmholtz 3D.
 * P2M, M2M, L2L and L2P are driven by the expansion tables alone: for the
 * Helmholtz kernel (real k) the caller passes one real plane of the complex
 * coefficients (and of the weights / outputs) at a time, with tables whose PDE
 * projection holds (k rscale)^2 for the expansion's rscale.
 * Precision: 0 = plain FP64; 1 = FP32 pair math with FP64 accumulation (P2P),
 * FP32 MFMA (M2L).
 */
#ifndef SUMPYHIP_H
#define SUMPYHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SUMPYHIP_OK 0
#define SUMPYHIP_EINVAL 1
#define SUMPYHIP_EHIP 2
#define SUMPYHIP_ENOTIMPL 3

/* Index tables of one expansion (kind, order), see sumpy_amd/expansion. */
typedef struct sumpyhip_expn {
  int order;
  int nfull;                    /* full Taylor identifiers |mi| <= order */
  int nstored;                  /* stored (PDE-compressed) identifiers */
  const int *full_ids;          /* 3 x nfull, axis-major */
  const int *stored_ids;        /* 3 x nstored */
  const int *stored_in_full;    /* nstored: full index of each stored id */
  const int *cube_index;        /* (order+1)^3: full index of (i,j,k), -1 if i+j+k>order */
  const int *proj_row_starts;   /* P (nfull x nstored) as CSR over full rows */
  const int *proj_cols;
  const double *proj_vals;
  const int *projT_col_starts;  /* P as CSC over stored columns */
  const int *projT_rows;
  const double *projT_vals;
  const int *horner_order;      /* nfull: full index at each Horner position */
  int proj_nnz;                 /* entries of P (proj_row_starts[nfull]) */
} sumpyhip_expn;

const char *sumpyhip_last_error(void);
int sumpyhip_device_info(int *n_cu, int *gcn_arch_major, int *gcn_arch_minor);

/* {{{ P2P (sumpy/p2p.py) */

/* Dense direct sum, replaces P2P.__call__ (sumpy/p2p.py:267-285):
 * result[t] = scale * sum_s [w0_s G(d) + w_s . grad_d G(d)], d = t - s,
 * plus the target gradient when gx/gy/gz are given.  target_to_source may be
 * NULL; otherwise the pair s == target_to_source[t] is skipped
 * (exclude_self, sumpy/p2p.py:162-170).  Overwrites pot/grad. */
int sumpyhip_p2p_dense(int kernel_id, int precision, double k_re, double k_im,
                       double scale, int ntargets, const double *tx, const double *ty,
                       const double *tz, int nsources, const double *sx,
                       const double *sy, const double *sz, const double *w0,
                       const double *wx, const double *wy, const double *wz,
                       const int *target_to_source, double *pot, double *gx,
                       double *gy, double *gz, void *stream);

/* Interaction matrix entries, replaces P2PMatrixGenerator.__call__ and
 * P2PMatrixSubsetGenerator.__call__ (sumpy/p2p.py:292-477): entry e is the
 * pair (tgtindices[e], srcindices[e]), or (e / nsources, e % nsources) for
 * the full row-major ntargets x nsources matrix (both index arrays NULL,
 * nentries = ntargets * nsources).  pot[e] = scale * the pair's kernel value
 * with the source's unit-strength weights w0 (and dipole weights wx/wy/wz,
 * NULL when no source derivative), the gradient outputs its target gradient
 * (NULL when not needed); a self pair (target_to_source[t] == s, when given)
 * is zero.  Complex (Helmholtz) weights and outputs are interleaved. */
int sumpyhip_p2p_matrix(int kernel_id, double k_re, double k_im, double scale,
                        int ntargets, const double *tx, const double *ty,
                        const double *tz, int nsources, const double *sx,
                        const double *sy, const double *sz, const void *w0,
                        const void *wx, const void *wy, const void *wz,
                        const int *target_to_source, long long nentries,
                        const int *tgtindices, const int *srcindices, void *pot,
                        void *gx, void *gy, void *gz, void *stream);

/* Near-field sum over per-box CSR source lists, replaces P2PFromCSR.__call__
 * (sumpy/p2p.py:801-833; loop structure :551-670).  Overwrites the results
 * of the targets in target_boxes.  box_centers (dim x nboxes) give the local
 * origin of the FP32 form; when box_centers is NULL the first target of each
 * target box is that origin (positions stay FP64 in the FP64 form). */
int sumpyhip_p2p_csr(int kernel_id, int precision, double k_re, double k_im,
                     double scale, int ntgt_boxes, const int *target_boxes,
                     const int *source_box_starts, const int *source_box_lists,
                     const int *box_source_starts, const int *box_source_counts_nonchild,
                     const int *box_target_starts, const int *box_target_counts_nonchild,
                     const double *box_centers, int nboxes, const double *tx,
                     const double *ty, const double *tz, const double *sx,
                     const double *sy, const double *sz, const double *w0,
                     const double *wx, const double *wy, const double *wz,
                     const int *target_to_source, double *pot, double *gx,
                     double *gy, double *gz, void *stream);

/* }}} */

/* {{{ expansions (Laplace 3D, Taylor / PDE-compressed Taylor) */

/* P2M, replaces P2EFromSingleBox.__call__ (sumpy/p2e.py:155-261):
 * tgt_expansions[b - tgt_base_ibox, :] = coefficients of the sources of box b
 * (overwrite).  wx/wy/wz may be NULL (no dipole sources). */
int sumpyhip_p2m(int kernel_id, const sumpyhip_expn *expn, int nsrc_boxes,
                 const int *source_boxes, const int *box_source_starts,
                 const int *box_source_counts_nonchild, const double *box_centers,
                 int nboxes, const double *sx, const double *sy, const double *sz,
                 const double *w0, const double *wx, const double *wy,
                 const double *wz, double rscale, double *tgt_expansions,
                 int tgt_base_ibox, void *stream);

/* P2L from list 4, replaces P2EFromCSR.__call__ (sumpy/p2e.py:268-388),
 * overwrite. */
int sumpyhip_p2l(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                 const int *target_boxes, const int *source_box_starts,
                 const int *source_box_lists, const int *box_source_starts,
                 const int *box_source_counts_nonchild, const double *box_centers,
                 int nboxes, const double *sx, const double *sy, const double *sz,
                 const double *w0, const double *wx, const double *wy,
                 const double *wz, double rscale, double *tgt_expansions,
                 int tgt_base_ibox, void *stream);

/* M2M, replaces E2EFromChildren.__call__ (sumpy/e2e.py:866-978):
 * tgt[parent] += sum_children T(c_parent - c_child) src[child]. */
int sumpyhip_m2m(int kernel_id, const sumpyhip_expn *src_expn,
                 const sumpyhip_expn *tgt_expn, int ntgt_boxes,
                 const int *target_boxes, const int *box_child_ids,
                 const double *box_centers, int nboxes, double src_rscale,
                 double tgt_rscale, const double *src_expansions, int src_base_ibox,
                 double *tgt_expansions, int tgt_base_ibox, void *stream);

/* L2L, replaces E2EFromParent.__call__ (sumpy/e2e.py:985-1083):
 * tgt[child] += T(c_child - c_parent) src[parent]. */
int sumpyhip_l2l(int kernel_id, const sumpyhip_expn *src_expn,
                 const sumpyhip_expn *tgt_expn, int ntgt_boxes,
                 const int *target_boxes, const int *box_parent_ids,
                 const double *box_centers, int nboxes, double src_rscale,
                 double tgt_rscale, const double *src_expansions, int src_base_ibox,
                 double *tgt_expansions, int tgt_base_ibox, void *stream);

/* Octree level translation as dense products, replaces E2EFromChildren.__call__
 * (M2M, sumpy/e2e.py:866-978) and E2EFromParent.__call__ (L2L,
 * sumpy/e2e.py:985-1083) when every child of the level sits at one of eight
 * fixed offsets from its parent: for each group of 16 target boxes and each
 * octant o with bit o set in group_octant_mask[g],
 *   tgt[b] += T_o src[group_inputs[g][o][j]]  (b = group_targets[g][j];
 *   -1 entries are skipped), accumulate.
 * T_frag holds the eight n_out x n_in FP64 matrices in MFMA A-fragment order:
 * T_frag[((o*nrb + rb)*ksteps + ks)*64 + l] = T_o[16 rb + l%16][4 ks + l/16]
 * (nrb = ceil(n_out/16), ksteps = ceil(n_in/4), zero padded).  A target box
 * may appear in one group only.  items: the nitems values g * 8 + o of the
 * octants present (bit o of group_octant_mask[g]), in group order.  partial:
 * workspace of ngroups * 8 * 16 * n_out doubles, through which the octants'
 * sums are added in octant order; NULL is allowed when every target has an
 * input in at most one octant (L2L). */
int sumpyhip_octant_translate(int ngroups, const int *group_targets,
                              const int *group_inputs, const int *group_octant_mask,
                              int nitems, const int *items, int n_out, int n_in,
                              const double *T_frag,
                              const double *src_expansions, int src_base_ibox,
                              double *tgt_expansions, int tgt_base_ibox, void *partial,
                              void *stream);

/* L2P, replaces E2PFromSingleBox.__call__ (sumpy/e2p.py:123-226); overwrites
 * pot (and the gradient when given) for the targets of each box. */
int sumpyhip_l2p(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                 const int *target_boxes, const int *box_target_starts,
                 const int *box_target_counts_nonchild, const double *box_centers,
                 int nboxes, const double *tx, const double *ty, const double *tz,
                 double rscale, double scale, const double *src_expansions,
                 int src_base_ibox, double *pot, double *gx, double *gy, double *gz,
                 void *stream);

/* M2P from list 3, replaces E2PFromCSR.__call__ (sumpy/e2p.py:233-343);
 * accumulates into pot (and the gradient when given). */
int sumpyhip_m2p(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                 const int *target_boxes, const int *source_box_starts,
                 const int *source_box_lists, const int *box_target_starts,
                 const int *box_target_counts_nonchild, const double *box_centers,
                 int nboxes, const double *tx, const double *ty, const double *tz,
                 double rscale, double scale, const double *src_expansions,
                 int src_base_ibox, double *pot, double *gx, double *gy, double *gz,
                 void *stream);

/* P2M for source boxes of several tree levels in one launch, replaces the
 * per-level loop of SumpyExpansionWrangler.form_multipoles
 * (sumpy/fmm.py:609-641) over P2EFromSingleBox.__call__ (sumpy/p2e.py:
 * 155-261) for the Laplace 3D kernel when every level has the same expansion
 * order: box b has the rscale level_rscale[box_levels[b]] and its multipole
 * at row b of tgt_expansions (one row per tree box); overwrites.  precision 1:
 * the source moments and their sums over a box in FP32 (orders 1-12 with the
 * full Taylor set), projected to the stored coefficients in FP64. */
int sumpyhip_p2m_levels(int kernel_id, const sumpyhip_expn *expn, int nsrc_boxes,
                        const int *source_boxes, const int *box_source_starts,
                        const int *box_source_counts_nonchild, const double *box_centers,
                        int nboxes, const double *sx, const double *sy, const double *sz,
                        const double *w0, const double *wx, const double *wy,
                        const double *wz, const int *box_levels, int nlevels,
                        const double *level_rscale, double *tgt_expansions, int precision,
                        void *stream);

/* L2P for target boxes of several tree levels in one launch, replaces the
 * per-level loop of SumpyExpansionWrangler.eval_locals (sumpy/fmm.py:
 * 1047-1083) over E2PFromSingleBox.__call__ (sumpy/e2p.py:123-226) for the
 * Laplace 3D kernel when every level has the same expansion order: box tb
 * has the rscale level_rscale[box_levels[tb]] and its local at row tb of
 * src_expansions (one row per tree box); overwrites.  precision 0: FP64;
 * 1: the Horner evaluation in FP32 (potentials, orders 1..12; L_full = P L
 * formed in FP64). */
int sumpyhip_l2p_levels(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                        const int *target_boxes, const int *box_target_starts,
                        const int *box_target_counts_nonchild, const double *box_centers,
                        int nboxes, const double *tx, const double *ty, const double *tz,
                        const int *box_levels, int nlevels, const double *level_rscale,
                        double scale, const double *src_expansions, double *pot, double *gx,
                        double *gy, double *gz, int precision, void *stream);

/* M2P from list 3 over source boxes of several tree levels in one launch,
 * replaces the per-level loop of SumpyExpansionWrangler.eval_multipoles
 * (sumpy/fmm.py:922-961) over E2PFromCSR.__call__ (sumpy/e2p.py:233-343)
 * when every level has the same expansion order: source box sb has the
 * rscale level_rscale[box_levels[sb]] (nlevels <= 32) and its multipole at
 * row sb of src_expansions (one row per tree box).  The source lists of a
 * target box are summed in list order; accumulates like sumpyhip_m2p.
 * precision 0: FP64; 1: each (target, source box) term in FP32 (recurrence,
 * multipole in FP32), the sums over source boxes in FP64. */
int sumpyhip_m2p_levels(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                        const int *target_boxes, const int *source_box_starts,
                        const int *source_box_lists, const int *box_target_starts,
                        const int *box_target_counts_nonchild, const double *box_centers,
                        int nboxes, const double *tx, const double *ty, const double *tz,
                        const int *box_levels, int nlevels, const double *level_rscale,
                        double scale, const double *src_expansions, double *pot, double *gx,
                        double *gy, double *gz, int precision, void *stream);

/* P2L from list 4 for target boxes of several tree levels in one launch,
 * replaces the per-level loop of SumpyExpansionWrangler.form_locals
 * (sumpy/fmm.py:963-1002) over P2EFromCSR.__call__ (sumpy/p2e.py:268-388)
 * when every level has the same expansion order: target box tb has the
 * rscale level_rscale[box_levels[tb]] and its local at row tb of
 * tgt_expansions (one row per tree box); overwrites like sumpyhip_p2l.
 * precision 0: FP64; 1: the derivatives and the sums over 64 sources in
 * FP32, the sums over those batches in FP64. */
int sumpyhip_p2l_levels(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                        const int *target_boxes, const int *source_box_starts,
                        const int *source_box_lists, const int *box_source_starts,
                        const int *box_source_counts_nonchild, const double *box_centers,
                        int nboxes, const double *sx, const double *sy, const double *sz,
                        const double *w0, const double *wx, const double *wy,
                        const double *wz, const int *box_levels, int nlevels,
                        const double *level_rscale, double *tgt_expansions, int precision,
                        void *stream);

/* }}} */

/* {{{ Helmholtz-3D (G = e^{ik|d|}/|d|, sumpy/kernel.py:754-844), k = k + i k_im
 * (k_im > 0: damped waves, full-Taylor expansions on the Python side).  M2P and
 * P2L take `precision` 0 (FP64) or 1 (FP32 terms with FP64 sums, the
 * polynomial and moment forms of the PDE-compressed expansions). */

/* M2P from list 3 for the Helmholtz kernel, replaces E2PFromCSR.__call__
 * (sumpy/e2p.py:233-343) for HelmholtzKernel(3); accumulates.  Multipoles,
 * pot and the gradient are complex, interleaved (re, im) doubles; expn holds
 * the stored identifiers of the PDE-compressed expansion (at most 256). */
int sumpyhip_m2p_helmholtz(const sumpyhip_expn *expn, double k, double k_im, int ntgt_boxes,
                           const int *target_boxes, const int *source_box_starts,
                           const int *source_box_lists, const int *box_target_starts,
                           const int *box_target_counts_nonchild,
                           const double *box_centers, int nboxes, const double *tx,
                           const double *ty, const double *tz, double rscale,
                           double scale, const void *src_expansions, int src_base_ibox,
                           void *pot, void *gx, void *gy, void *gz, int precision, void *stream);

/* P2L from list 4 for the Helmholtz kernel, replaces P2EFromCSR.__call__
 * (sumpy/p2e.py:268-388) for HelmholtzKernel(3); overwrites the rows of the
 * target boxes.  Weights and expansions are complex, interleaved. */
int sumpyhip_p2l_helmholtz(const sumpyhip_expn *expn, double k, double k_im, int ntgt_boxes,
                           const int *target_boxes, const int *source_box_starts,
                           const int *source_box_lists, const int *box_source_starts,
                           const int *box_source_counts_nonchild,
                           const double *box_centers, int nboxes, const double *sx,
                           const double *sy, const double *sz, const void *w0,
                           const void *wx, const void *wy, const void *wz, double rscale,
                           void *tgt_expansions, int tgt_base_ibox, int precision, void *stream);

/* Class data for the Helmholtz kernel, replaces
 * M2LGenerateTranslationClassesDependentData.__call__ (sumpy/e2e.py:565-669)
 * for HelmholtzKernel(3): out is nclasses x 2 x ncirc doubles, the real then
 * the imaginary plane of rscale^|m| d^m G(v_c) in the circulant layout of
 * sumpyhip_m2l_class_data.  The two planes are the classes 2c and 2c + 1 of
 * the real embedding that sumpyhip_m2l_apply contracts. */
int sumpyhip_m2l_class_data_helmholtz(int order_src, int order_tgt, const int *circ_shape,
                                      int nclasses, const double *translation_vectors,
                                      int nvectors, int class_start, double src_rscale,
                                      double k, double k_im, double *out, void *stream);

/* M2L without translation classes for the Helmholtz kernel, replaces
 * E2EFromCSR.__call__ (sumpy/e2e.py:164-304) for HelmholtzKernel(3): the
 * arguments of sumpyhip_m2l_csr plus the wavenumber k + i k_im; complex
 * expansions, interleaved; overwrites the target rows. */
int sumpyhip_m2l_csr_helmholtz(int order_src, int order_tgt, const int *circ_shape, int nl,
                               int nm, const int *off_t, const int *off_s,
                               const int *tgt_degree, int ntgt_boxes, const int *target_boxes,
                               const int *src_box_starts, const int *src_box_lists,
                               const double *box_centers, int nboxes, double src_rscale,
                               double tgt_rscale, double k, double k_im,
                               const void *src_expansions,
                               int src_base_ibox, void *tgt_expansions, int tgt_base_ibox,
                               void *stream);

/* }}} */

/* {{{ M2L through translation classes */

/* Class data, replaces M2LGenerateTranslationClassesDependentData.__call__
 * (sumpy/e2e.py:565-669): for class c, out[c, circ(m)] = rscale^|m| d^m G(v_c)
 * for the derivatives m = s + t the translation needs, in the circulant layout
 * of the reference's non-FFT M2L (sumpy/expansion/m2l.py:403-537); other
 * entries are zero.  circ_shape = max_mi + 1 (3 ints). */
int sumpyhip_m2l_class_data(int kernel_id, int order_src, int order_tgt,
                            const int *circ_shape, int nclasses,
                            const double *translation_vectors, int nvectors,
                            int class_start, double src_rscale, double *out,
                            void *stream);

/* Preprocessing, replaces M2LPreprocessMultipole.__call__
 * (sumpy/e2e.py:676-760): pack stored multipoles into rows of kpad entries of
 * the M2L compute precision (precision 0: FP64, 1: FP32, 2: kpad bf16 high
 * parts then kpad bf16 low parts).  Entry k of a row holds stored coefficient
 * k_source[k] (zero where -1); k_source NULL: entry k holds coefficient k
 * (zero past nstored). */
int sumpyhip_m2l_preprocess(int precision, int nboxes, int nstored,
                            const double *src_expansions, int kpad, const int *k_source,
                            void *packed, void *stream);

/* One tree level of sumpyhip_m2l_apply: tile schedule (built on the host),
 * class data (nclasses x ncirc: FP64, or at precision 2 the packed words of
 * sumpyhip_m2l_pack_bf16_class_data), packed source multipoles (compute
 * precision, kpad per box), work rows (compute precision, npad per box) and,
 * when nsplit > 1, a partial buffer of (nsplit - 1) * ntiles * 64 * npad
 * entries of the compute precision. */
typedef struct sumpyhip_m2l_level {
  int ntiles, nsplit;
  const int *tile_targets;      /* ntiles x tile_width target boxes (-1: none) */
  const int *tile_slot_starts;  /* ntiles + 1 */
  const int *slot_class;        /* nslots: class of the slot (level-relative) */
  const int *slot_sources;      /* nslots x tile_width source boxes (-1: none) */
  const void *class_data;
  const void *packed_src;
  int src_base_ibox;
  void *work;
  int tgt_base_ibox;
  void *partial;
  int tile_width;               /* targets per tile: 64 (0 means 64) or 32; the
                                   levels of one call share it */
} sumpyhip_m2l_level;

/* Translation, replaces M2LUsingTranslationClassesDependentData.__call__
 * (sumpy/e2e.py:310-558) as a class-batched dense contraction on MFMA, for up
 * to 32 levels of the same orders in one launch: for each tile of 64 (or 32,
 * the levels' tile_width) target boxes and each class slot of the tile,
 * Y += A_c X with A_c[t, s] = data_c[off_t[t] + off_s[s]] and X the stacked
 * packed multipoles of the tile's sources.  Overwrites the work rows of the
 * tile targets.  The slots of a tile are split over nsplit blocks; splits
 * 1.. write partial sums that are then added to the work rows in a fixed
 * order (deterministic). */
int sumpyhip_m2l_apply(int precision, int nlevels, const sumpyhip_m2l_level *levels,
                       int nl, int nm, int npad, int kpad, const int *off_t,
                       const int *off_s, int ncirc, void *stream);

/* Bytes of the fixed device scratch of sumpyhip_m2l_apply_noalloc. */
#define SUMPYHIP_M2L_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sumpyhip_m2l_apply with the same arguments and results,
 * except that levels[l].partial is ignored -- the split partial sums of all
 * levels live in a fixed device scratch of SUMPYHIP_M2L_NOALLOC_SCRATCH_BYTES
 * per GPU (held until the process ends), and a call whose partial sums need
 * more (sum over levels with nsplit > 1 of (nsplit - 1) * ntiles * 64 * npad
 * values of the compute precision, each level rounded up to 256 bytes) fails
 * with an argument error.  Allocates nothing.  Calls that overlap on one GPU
 * (other streams or host threads) race on the scratch. */
int sumpyhip_m2l_apply_noalloc(int precision, int nlevels, const sumpyhip_m2l_level *levels,
                               int nl, int nm, int npad, int kpad, const int *off_t,
                               const int *off_s, int ncirc, void *stream);

/* Split-BF16 form (precision 2) support, no reference counterpart (a layout
 * of this library's split-BF16 contraction): the class data of nclasses
 * classes (circulant shape circ_shape, as written by sumpyhip_m2l_class_data)
 * as packed words bf16(x) | bf16(x - bf16(x)) << 16, the value of multi-index
 * (mx, my, mz) at word mx S0 + my S1 + mz of a class's ncirc_packed words
 * (zero in the gaps).  At precision 2, sumpyhip_m2l_apply takes these words
 * as class_data, with off_t / off_s computed with the same strides: off_t has
 * npad entries (pass nl = npad) and off_s kpad entries (pass nm = kpad), and
 * ncirc = ncirc_packed.  Row and k orders are free (sumpyhip_m2l_preprocess's
 * k_source, sumpyhip_m2l_postprocess's row_pos). */
int sumpyhip_m2l_pack_bf16_class_data(int nclasses, const int *circ_shape, int S0, int S1,
                                      int ncirc_packed, const double *class_data,
                                      void *packed, void *stream);

/* Postprocessing, replaces M2LPostprocessLocal.__call__
 * (sumpy/e2e.py:767-859): locals[b, i] = work[b, row_pos[i]] * ratio^deg_i
 * (row_pos NULL: row i). */
int sumpyhip_m2l_postprocess(int precision, int nboxes, int nl, int npad,
                             const void *work, const int *tgt_degree, const int *row_pos,
                             double ratio, double *tgt_expansions, void *stream);

/* M2L without translation classes, replaces E2EFromCSR.__call__ for
 * multipole-to-local (sumpy/e2e.py:164-304); overwrites. */
int sumpyhip_m2l_csr(int kernel_id, int order_src, int order_tgt,
                     const int *circ_shape, int nl, int nm, const int *off_t,
                     const int *off_s, const int *tgt_degree, int ntgt_boxes,
                     const int *target_boxes, const int *src_box_starts,
                     const int *src_box_lists, const double *box_centers, int nboxes,
                     double src_rscale, double tgt_rscale, const double *src_expansions,
                     int src_base_ibox, double *tgt_expansions, int tgt_base_ibox,
                     void *stream);

/* }}} */

/* {{{ source weights and target outputs */

/* w[a][s] += coeff * strength[s] * (dir ? dir[a][s] : 1): builds the charge
 * (a = -1 -> w0) and dipole arrays from the reference's source kernels
 * (AxisSourceDerivative / DirectionalSourceDerivative, sumpy/kernel.py:2288-2507).
 * Complex when is_complex (interleaved). */
int sumpyhip_axpy_weights(int is_complex, int n, double coeff, const double *strength,
                          const double *dir, double *w, void *stream);

/* out[t] = sum_j coeff_j * (mult_j ? x_{mult_j}[t] : 1) * in_j[t]: combines
 * potential/gradient into TargetPointMultiplier / AxisTargetDerivative results
 * (sumpy/kernel.py:2330-2588). */
int sumpyhip_combine_outputs(int is_complex, int n, int nterms, const double *const *ins,
                             const double *coeffs, const double *const *mults,
                             double *out, int accumulate, void *stream);

/* }}} */

/* {{{ QBX */

/* Layer-potential matrix entries, replaces LayerPotentialMatrixGenerator.__call__
 * and LayerPotentialMatrixSubsetGenerator.__call__ (sumpy/qbx.py:342-539):
 * entry e of the pair (tgtindices[e], srcindices[e]) (or of the full
 * row-major matrix, index arrays NULL) is scale times the term of
 * sumpyhip_qbx2d for a unit density; accumulate = 1 adds to out (one call per
 * source kernel). */
int sumpyhip_qbx2d_matrix(int order, int src_kind, int tgt_axis, int ntargets,
                          const double *tx, const double *ty, const double *cx,
                          const double *cy, int nsources, const double *sx,
                          const double *sy, const double *nx, const double *ny,
                          long long nentries, const int *tgtindices, const int *srcindices,
                          double scale, int accumulate, double *out, void *stream);

/* 2D Laplace line-Taylor layer potential, replaces LayerPotential.__call__
 * (sumpy/qbx.py:250-335) for LineTaylorLocalExpansion(LaplaceKernel(2)):
 * out[t] = scale * sum_s sigma_s Re sum_{i<=order} F^(i)(a) b^i / i!,
 * a = c_t - s, b = t - c_t (complex notation), F = log (src_kind 0, SLP),
 * -n/a (1, DLP), -e_axis/a (2+axis, AxisSourceDerivative); tgt_axis >= 0
 * applies d/da_axis.  nx/ny: source normals (DLP). */
int sumpyhip_qbx2d(int order, int src_kind, int tgt_axis, int ntargets,
                   const double *tx, const double *ty, const double *cx,
                   const double *cy, int nsources, const double *sx,
                   const double *sy, const double *nx, const double *ny,
                   const double *density, double scale, double *out, void *stream);

/* Line-Taylor layer potential for the Laplace, Helmholtz and Yukawa kernels in
 * 2D and 3D, replaces LayerPotential.__call__ (sumpy/qbx.py:250-335) with
 * LineTaylorLocalExpansion (sumpy/expansion/local.py:116-198) for those
 * kernels: result[t] (complex, re/im interleaved) = scale * sum_s sigma_s
 * sum_{i<=order} [tau^i] D_s K(a + tau b), a = c_t - s, b = t - c_t.
 * kernel_kind: 0 Laplace 2D (log r), 1 Laplace 3D (1/r), 2 Helmholtz 2D
 * (H0(k r), k = k_re + i k_im, k_im >= 0), 3 Helmholtz 3D (e^{ikr}/r),
 * 4 Yukawa 2D (H0(i lam r)), 5 Yukawa 3D (e^{-lam r}/r); lam = k_re.
 * src_kind: 0 none, 1 -n.grad_a (DirectionalSourceDerivative, normals
 * nx/ny/nz), 2 -d/da_src_axis (AxisSourceDerivative); tgt_axis >= 0 applies
 * d/da_tgt_axis (AxisTargetDerivative).  tz, cz, sz, nz: NULL in 2D.
 * order <= 12.  accumulate = 1 adds to result. */
int sumpyhip_qbx_line(int kernel_kind, int order, int src_kind, int src_axis, int tgt_axis,
                      double k_re, double k_im, int ntargets, const double *tx,
                      const double *ty, const double *tz, const double *cx, const double *cy,
                      const double *cz, int nsources, const double *sx, const double *sy,
                      const double *sz, const double *nx, const double *ny, const double *nz,
                      const double *density, double scale_re, double scale_im, int accumulate,
                      double *result, void *stream);

/* Matrix entries of sumpyhip_qbx_line (unit density), replaces
 * LayerPotentialMatrixGenerator / LayerPotentialMatrixSubsetGenerator
 * (sumpy/qbx.py:342-539) for the same kernels: entry e of the pair
 * (tgt_idx[e], src_idx[e]), or of the full row-major matrix (index arrays
 * NULL, nentries = ntargets * nsources); complex, re/im interleaved. */
int sumpyhip_qbx_line_matrix(int kernel_kind, int order, int src_kind, int src_axis,
                             int tgt_axis, double k_re, double k_im, int ntargets,
                             const double *tx, const double *ty, const double *tz,
                             const double *cx, const double *cy, const double *cz,
                             int nsources, const double *sx, const double *sy, const double *sz,
                             const double *nx, const double *ny, const double *nz,
                             long long nentries, const int *tgt_idx, const int *src_idx,
                             double scale_re, double scale_im, int accumulate, double *out,
                             void *stream);

/* Adjoint of sumpyhip_qbx2d for the single (src_kind 0) and double (1) layer
 * without target derivatives: with g = grad_out (per target), the gradients of
 * sum_t g_t out_t with respect to density, source positions, source normals
 * (DLP), target positions and expansion centres.  Any output pair may be
 * NULL.  Two kernels: one parallel over targets (d/dt, d/dc), one over
 * sources (d/dsigma, d/ds, d/dn).  No reference counterpart (sumpy has no
 * adjoints); the math is the a- and b-derivative of the line-Taylor sum of
 * sumpy/expansion/local.py:132-185. */
int sumpyhip_qbx2d_adjoint(int order, int src_kind, int ntargets, const double *tx,
                           const double *ty, const double *cx, const double *cy,
                           int nsources, const double *sx, const double *sy,
                           const double *nx, const double *ny, const double *density,
                           double scale, const double *grad_out, double *d_density,
                           double *d_sx, double *d_sy, double *d_nx, double *d_ny,
                           double *d_tx, double *d_ty, double *d_cx, double *d_cy,
                           void *stream);

/* }}} */

/* {{{ 2D volume-Taylor expansions: Laplace (kind 0, G = log r, real rows) and
 * Helmholtz (kind 1, G = H0(k r), k = k + i k_im with k_im >= 0, complex
 * rows; Laplace takes k_im = 0), full and
 * PDE-compressed (VolumeTaylor* / LinearPDEConformingVolumeTaylor*
 * expansions of sumpy/expansion/multipole.py:70-419, local.py:200-510 in
 * 2D).  Expansions carry no scaling constant; M2P and L2P multiply by
 * scale_re + i scale_im.  Orders 0..12. */

/* One expansion: its order, whether it is PDE-compressed (stored m_x <= 1,
 * the rest by u_{mx,my} = -u_{mx-2,my+2} - (k rscale)^2 u_{mx-2,my}) and its
 * stored multi-indices in storage order (device arrays of nstored ints). */
typedef struct sumpyhip_t2d_expn {
  int order, compressed, nstored;
  const int *stored_mx, *stored_my;
} sumpyhip_t2d_expn;

/* P2M, replaces P2EFromSingleBox.__call__ (sumpy/p2e.py:155-261): overwrites
 * the rows of the source boxes (row box - tgt_base_ibox).  weights_dx /
 * weights_dy (NULL: none): dipole weights u contributing u . grad_d G(d),
 * d = t - s, as in the H2D functions. */
int sumpyhip_t2d_p2m(int kind, const sumpyhip_t2d_expn *ex, double k, double k_im, int nboxes,
                     const int *source_boxes, const int *box_source_starts,
                     const int *box_source_counts, const double *centers, int ncenters,
                     const double *sx, const double *sy, const double *weights,
                     const double *weights_dx, const double *weights_dy, double rscale,
                     double *tgt_expansions, int tgt_base_ibox, void *stream);

/* L2P, replaces E2PFromSingleBox.__call__ (sumpy/e2p.py:123-226): writes the
 * potentials (result) and d/dx, d/dy (result_dx, result_dy) of the target
 * boxes' targets; any of the three may be NULL. */
int sumpyhip_t2d_l2p(int kind, const sumpyhip_t2d_expn *ex, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *box_target_starts,
                     const int *box_target_counts, const double *centers, int ncenters,
                     const double *tx, const double *ty, double rscale,
                     const double *src_expansions, int src_base_ibox, double scale_re,
                     double scale_im, double *result, double *result_dx, double *result_dy,
                     void *stream);

/* M2P, replaces E2PFromCSR.__call__ (sumpy/e2p.py:233-343): adds the list-3
 * multipoles' potentials (and gradients, as in L2P). */
int sumpyhip_t2d_m2p(int kind, const sumpyhip_t2d_expn *ex, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *source_box_starts,
                     const int *source_box_lists, const int *box_target_starts,
                     const int *box_target_counts, const double *centers, int ncenters,
                     const double *tx, const double *ty, double rscale,
                     const double *src_expansions, int src_base_ibox, double scale_re,
                     double scale_im, double *result, double *result_dx, double *result_dy,
                     void *stream);

/* P2L, replaces P2EFromCSR.__call__ (sumpy/p2e.py:268-388): overwrites the
 * target boxes' rows from their list-4 sources (dipole weights as in P2M). */
int sumpyhip_t2d_p2l(int kind, const sumpyhip_t2d_expn *ex, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *source_box_starts,
                     const int *source_box_lists, const int *box_source_starts,
                     const int *box_source_counts, const double *centers, int ncenters,
                     const double *sx, const double *sy, const double *weights,
                     const double *weights_dx, const double *weights_dy, double rscale,
                     double *tgt_expansions, int tgt_base_ibox, void *stream);

/* M2M, replaces E2EFromChildren.__call__ (sumpy/e2e.py:866-978): adds the
 * translated children (box_child_ids: 4 x child_stride, 0 = none). */
int sumpyhip_t2d_m2m(int kind, const sumpyhip_t2d_expn *src_ex, const sumpyhip_t2d_expn *tgt_ex,
                     double k, double k_im, int ntgt, const int *target_boxes, const int *box_child_ids,
                     int child_stride, const double *centers, int ncenters, double src_rscale,
                     double tgt_rscale, const double *src_expansions, int src_base_ibox,
                     double *tgt_expansions, int tgt_base_ibox, void *stream);

/* L2L, replaces E2EFromParent.__call__ (sumpy/e2e.py:985-1090): adds the
 * translated parent local. */
int sumpyhip_t2d_l2l(int kind, const sumpyhip_t2d_expn *src_ex, const sumpyhip_t2d_expn *tgt_ex,
                     double k, double k_im, int ntgt, const int *target_boxes, const int *box_parent_ids,
                     const double *centers, int ncenters, double src_rscale, double tgt_rscale,
                     const double *src_expansions, int src_base_ibox, double *tgt_expansions,
                     int tgt_base_ibox, void *stream);

/* M2L class data, replaces M2LGenerateTranslationClassesDependentData
 * (sumpy/e2e.py:565-669): row c holds rscale^|m| d^m G(vec[class_start + c])
 * for |m| <= order_sum at index |m|(|m|+1)/2 + m_y (complex: re, im). */
int sumpyhip_t2d_m2l_data(int kind, int order_sum, double k, double k_im, int ncls, const double *vecs,
                          int nvecs, int class_start, double rscale, double *out, void *stream);

/* M2L, replaces M2LUsingTranslationClassesDependentData.__call__ (class data
 * given, the class of list entry e in classes[e]) and E2EFromCSR.__call__
 * (class data NULL; derivatives per pair) (sumpy/e2e.py:164-558): overwrites
 * each target box's row. */
int sumpyhip_t2d_m2l(int kind, const sumpyhip_t2d_expn *src_ex, const sumpyhip_t2d_expn *tgt_ex,
                     double k, double k_im, int ntgt, const int *target_boxes, const int *src_box_starts,
                     const int *src_box_lists, const int *classes, int class_start,
                     const double *class_data, const double *centers, int ncenters,
                     double src_rscale, double tgt_rscale, const double *src_expansions,
                     int src_base_ibox, double *tgt_expansions, int tgt_base_ibox,
                     void *stream);

/* }}} */

/* {{{ 2D Hankel/Bessel expansions: Helmholtz (H2D, kind 0, G = (i/4) H0^(1)(k|d|))
 * and Yukawa (Y2D, kind 1, G = (i/4) H0^(1)(i lam |d|) = K0(lam |d|)/(2 pi))

 * The reference's H2D / Y2D multipole and local expansions
 * (sumpy/expansion/multipole.py:426-567, sumpy/expansion/local.py:522-704)
 * with the Fourier-Bessel M2L (sumpy/expansion/m2l.py:1008-1115), behind the
 * same operators as the 3D expansions (P2EFromSingleBox, P2EFromCSR,
 * E2EFromChildren, E2EFromParent, E2EFromCSR /
 * M2LUsingTranslationClassesDependentData, E2PFromSingleBox, E2PFromCSR,
 * P2PFromCSR, P2P: sumpy/p2e.py, e2e.py, e2p.py, p2p.py).  `k` + i `k_im` is the
 * Helmholtz wavenumber (k_im >= 0: damped waves; k_im = 0: real k > 0, which
 * takes the real-argument Bessel route) or `k` the Yukawa lambda (real, > 0,
 * k_im = 0), orders 0..31, FP64.  Expansions, weights and outputs are complex128 (interleaved
 * re, im); an expansion row holds the 2p+1 coefficients c = -p..p at c + p.
 * Coordinates are SoA; centers is 2 x ncenters.  The evaluations include the
 * kernel's i/4; M2P adds to `out`, L2P and the P2P forms write it.  Source
 * dipoles (DirectionalSourceDerivative) and target gradients
 * (AxisTargetDerivative) use the shifts W Z_c(s r) e^{icth} =
 * -s Z_{c+1} e^{i(c+1)th}, Wbar = s Z_{c-1} e^{i(c-1)th} (W = d_x + i d_y,
 * s = k or i lam). */

/* P2M of each source box from its own sources (rows box - base of out).
 * weights: charges; weights_dx / weights_dy (either may be NULL): dipole
 * weights u contributing u . grad_d G(d), d = t - s (as in the 3D P2P; a
 * DirectionalSourceDerivative source of strength q along n has u = -q n). */
int sumpyhip_h2d_p2m(int kind, int order, double k, double k_im, int nboxes,
                     const int *source_boxes, const int *box_source_starts,
                     const int *box_source_counts, const double *centers, int ncenters,
                     const double *sx, const double *sy, const void *weights,
                     const void *weights_dx, const void *weights_dy, double rscale, void *out,
                     int base, void *stream);

/* P2L of each target box from the sources of its list-4 boxes (CSR); weights
 * as in P2M. */
int sumpyhip_h2d_p2l(int kind, int order, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *starts, const int *lists,
                     const int *box_source_starts, const int *box_source_counts,
                     const double *centers, int ncenters, const double *sx, const double *sy,
                     const void *weights, const void *weights_dx, const void *weights_dy,
                     double rscale, void *out, int base, void *stream);

/* M2M: each target box's children (box_child_ids, 4 x nboxes, 0 = none)
 * translated and added to its row. */
int sumpyhip_h2d_m2m(int kind, int src_order, int tgt_order, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *box_child_ids, int nboxes,
                     const double *centers, int ncenters, double src_rscale,
                     double tgt_rscale, const void *src, int src_base, void *tgt,
                     int tgt_base, void *stream);

/* L2L: each target box's parent local translated and added to its row. */
int sumpyhip_h2d_l2l(int kind, int src_order, int tgt_order, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *box_parent_ids,
                     const double *centers, int ncenters, double src_rscale,
                     double tgt_rscale, const void *src, int src_base, void *tgt,
                     int tgt_base, void *stream);

/* Translation-class data of ncls classes (distance vectors vecs, 2 x nvecs,
 * from column cls_start): row c holds H_n(k|d|) e^{i n th_d},
 * n = -(src+tgt)..(src+tgt), as complex128. */
int sumpyhip_h2d_m2l_data(int kind, int src_order, int tgt_order, double k, double k_im, int ncls,
                          const double *vecs, int nvecs, int cls_start, void *data,
                          void *stream);

/* M2L over list 2 (CSR); with `data` (rows by classes[e] - cls_start) or,
 * when data is NULL, the Hankel row of each pair computed in place.  Writes
 * each target box's row. */
int sumpyhip_h2d_m2l(int kind, int src_order, int tgt_order, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *starts, const int *lists,
                     const int *classes, int cls_start, const void *data,
                     const double *centers, int ncenters, double src_rscale,
                     double tgt_rscale, const void *src, int src_base, void *tgt,
                     int tgt_base, void *stream);

/* L2P of each target box at its own targets (writes): the potential into out,
 * d/dx and d/dy (AxisTargetDerivative) into out_dx, out_dy; any of the three
 * may be NULL. */
int sumpyhip_h2d_l2p(int kind, int order, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *box_target_starts,
                     const int *box_target_counts, const double *centers, int ncenters,
                     const double *tx, const double *ty, double rscale, const void *loc,
                     int base, void *out, void *out_dx, void *out_dy, void *stream);

/* M2P from the list-3 boxes of each target box (adds); outputs as in L2P. */
int sumpyhip_h2d_m2p(int kind, int order, double k, double k_im, int ntgt,
                     const int *target_boxes, const int *starts, const int *lists,
                     const int *box_target_starts, const int *box_target_counts,
                     const double *centers, int ncenters, const double *tx, const double *ty,
                     double rscale, const void *mpoles, int base, void *out, void *out_dx,
                     void *out_dy, void *stream);

/* Near field over list 1 (CSR); t2s (or NULL): the source index of each
 * target to skip; pairs at zero distance add nothing.  Weights as in P2M,
 * outputs as in L2P (written). */
int sumpyhip_h2d_p2p_csr(int kind, double k, double k_im, int ntgt, const int *target_boxes,
                         const int *starts, const int *lists, const int *box_source_starts,
                         const int *box_source_counts, const int *box_target_starts,
                         const int *box_target_counts, const double *tx, const double *ty,
                         const double *sx, const double *sy, const void *weights,
                         const void *weights_dx, const void *weights_dy, const int *t2s,
                         void *out, void *out_dx, void *out_dy, void *stream);

/* Dense direct sum; weights and outputs as in the CSR form. */
int sumpyhip_h2d_p2p_dense(int kind, double k, double k_im, int ntargets, const double *tx,
                           const double *ty, int nsources, const double *sx, const double *sy,
                           const void *weights, const void *weights_dx, const void *weights_dy,
                           const int *t2s, void *out, void *out_dx, void *out_dy,
                           void *stream);

/* }}} */

/* {{{ tree and traversal on the GPU

 * Building blocks of the device tree / traversal builder
 * (sumpy_amd/tree_gpu.py orchestrates them: sorts, scans, compaction).  The
 * reference takes its tree, lists and translation classes from boxtree
 * (TreeBuilder, FMMTraversalBuilder, TranslationClassesBuilder, called at
 * sumpy/test/test_fmm.py:182-196 and read at sumpy/fmm.py:357-592); these
 * produce the same arrays as the host builder, array for array.  Integer
 * arrays are int32, keys and Morton prefixes uint64, coordinates FP64 SoA. */

/* Morton keys of n points (x, y, z: device arrays; z unused in 2D): cell
 * c_d = floor((x_d - lo[d]) / ext * 2^maxdepth) clamped to the grid, axis d
 * bit b at key bit dim*b + dim-1-d.  lo: host array of dim doubles. */
int sumpytree_gpu_morton(int dim, long n, const double *x, const double *y, const double *z,
                         const double *lo, double ext, int maxdepth, void *keys, void *stream);

/* Children of nsplit split boxes (Morton prefix, sorted source and target
 * ranges): for child c, the ranges of prefix (P << dim) | c among the sorted
 * keys (key >> shift), and whether it holds any particle.  tkeys NULL: the
 * targets are the sources.  Outputs have nsplit << dim entries. */
int sumpytree_gpu_children(int dim, int nsplit, const void *prefix, const int *slo,
                           const int *shi, const int *tlo, const int *thi, const void *skeys,
                           const void *tkeys, int shift, int *cslo, int *cshi, int *ctlo,
                           int *cthi, int *exists, void *stream);

/* Box centres (dim x nboxes, lo + (ix + 1/2) ext / 2^level) and integer cell
 * coordinates (dim x nboxes) from the boxes' Morton prefixes and levels. */
int sumpytree_gpu_geometry(int dim, int nboxes, const void *prefix, const int *level,
                           const double *lo, double ext, double *centers, int *ix,
                           void *stream);

/* Same-level neighbours of every box, ascending, -1 padded to 26 (3D) or 8
 * (2D) per box. */
int sumpytree_gpu_colleagues(int dim, int nboxes, int nlevels, const int *level,
                             const int *ix, const int *child, const int *parent,
                             const int *leaf, const int *level_start, const int *src_cumul,
                             const int *src_nonchild, const void *prefix, int *coll,
                             void *stream);

/* Interaction list `which` (1: neighbour leaves, 2: well-separated children of
 * the parent's colleagues, 3: separated smaller boxes per source level, 4:
 * separated coarser leaves) of nb boxes.  counts != NULL: the counting pass
 * (list 3: nlevels counts per box); else the writing pass at starts (list 3:
 * offsets per (level, box)), with the translation slot of each list-2 pair in
 * slots and the level's used slots flagged in used. */
int sumpytree_gpu_list(int which, int dim, int nboxes, int nlevels, const int *level,
                       const int *ix, const int *child, const int *parent, const int *leaf,
                       const int *level_start, const int *src_cumul, const int *src_nonchild,
                       const void *prefix, const int *coll, int nb, const int *boxes,
                       int *counts, const int *starts, int *out, int *slots, int *used,
                       void *stream);

/* Translation class of each of n list-2 pairs:
 * cls[e] = slot_to_cls[entry_level[e] * nslots + slot[e]]. */
int sumpytree_gpu_list2_classes(int n, const int *entry_level, const int *slot,
                                const int *slot_to_cls, int nslots, int *cls, void *stream);

/* }}} */

#ifdef __cplusplus
}
#endif

#endif
