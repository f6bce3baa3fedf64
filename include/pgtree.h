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

/* Periodic unit box: ghosts within rcut of a face are appended; i-groups are
 * octree nodes of <= ncrit particles; j-list = particles within rcut of the
 * group's bounding box. */
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

#if defined(__cplusplus)
}
#endif

#endif
