/* Host (CPU, OpenMP) interaction-list builder of libpgtree.so.
 * pgamd-library: pgtree
 *
 * Caller-side code that produces the inputs of the batched GPU calls.  The
 * reference ships its Barnes-Hut list builder inside the sample program
 * PG5/newton_quad/source.cpp (moments :215-354, opening test and walk
 * :581-707) and no TreePM list builder; these are this library's own.
 */
#ifndef PGTREE_H
#define PGTREE_H

#if defined(__cplusplus)
extern "C" {
#endif

/* Every particle belongs to exactly one i-group.  An i-group is an octree
 * node of <= ncrit particles whose parent holds more, or a node the 21-level
 * Morton key cannot split (more than ncrit particles inside one cell of side
 * L / 2^21, e.g. coincident particles), which is one group of all of them.
 * In a Barnes-Hut walk such a node, when opened, lists its particles as a
 * leaf does, so every mass is reachable.
 * Periodic unit box: ghosts within rcut of a face are appended; i-groups as
 * above; j-list = particles within rcut of the
 * group's bounding box, in the order of a uniform grid of cells of side
 * >= rcut over [-rcut, 1 + rcut) (cells by x, y, z index; a cell's particles
 * by index). */
void *pgt_treepm_build(int n, const double *pos, double rcut, int ncrit);
/* Open domain Barnes-Hut lists with the opening test of source.cpp:581-707. */
void *pgt_bh_build(int n, const double *pos, const double *mass, double theta, int ncrit,
                   int nleaf);
/* info[8] = {np, ngroups, len(ilist), len(jlist), len(clist), ncells,
 *            sum ni*nj, sum ni*ncells} */
int pgt_info(void *h, long long *info);
/* what: 0 pos, 1 src, 2 ioff, 3 icnt, 4 ilist, 5 joff, 6 jcnt, 7 jlist,
 * 8 coff, 9 ccnt, 10 clist, 11 order, 12 cells */
int pgt_get(void *h, int what, void *out);
void pgt_free(void *h);

/* ------------------------------------------------------------------ libpgtree_dev.so
 * pgamd-library: pgtree_dev
 * The same lists built on the GPU from positions already in HBM (all
 * pointers are device pointers; `stream` is a hipStream_t, null = the
 * library's own).  Array for array identical to the host builder's, except
 * the numbering of Barnes-Hut cells (the lists hold the same cell contents
 * in the same order).  No counterpart in the reference: its tree code builds
 * its lists on the CPU (PG5/newton_quad/source.cpp:215-707). */
void *pgt_dev_treepm_build(int n, const double *pos, double rcut, int ncrit, void *stream);
void *pgt_dev_bh_build(int n, const double *pos, const double *mass, double theta, int ncrit,
                       int nleaf, void *stream);
/* info[8] as pgt_info. */
int pgt_dev_info(void *h, long long *info);
/* Device pointer of array `what` (codes of pgt_get), valid until pgt_dev_free. */
const void *pgt_dev_ptr(void *h, int what);
/* Copies array `what` to host memory (sizes as pgt_get). */
int pgt_dev_get(void *h, int what, void *out);
/* Batch-call inputs on the device: pos4[np] = float4 {float(x) * float(scale),
 * ..., float(mass[src])}; cells12[ncells] = {x, y, z, M, Qxx, Qxy, Qxz, Qyy,
 * Qyz, Qzz, 0, 0} as floats.  Either output may be null. */
int pgt_dev_pack(void *h, const double *mass, double scale, void *pos4, void *cells12,
                 void *stream);
void pgt_dev_free(void *h);
void pgamd_set_abort_on_error(int flag);
int pgamd_last_error(char *buf, int buflen);

#if defined(__cplusplus)
}
#endif

#endif
