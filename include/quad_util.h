/* quad_tree (monopole + quadrupole cell) API of libpg5.so.
 * pgamd-library: pg5 pg5_sym
 *
 * Same declarations as PG5/newton_quad/quad/quad_util.h:15-35.  Cells carry
 * the centre of mass xj, mass mj and the traceless quadrupole
 * qj = {xx, xy, xz, yy, yz, zz} of sum m (3 d d - |d|^2 I)
 * (PG5/newton_quad/source.cpp:215-248).  The i-particles and softening come
 * from the GRAPE-5 state (g5_set_xiMC, g5_set_eps_to_all), as in quad_g5c.c.
 * Output: a = sum (phi_p + 5 phi_q) d / r^2 - Q d / r^5 and a NEGATIVE
 * p = -sum (M/r + d.Q.d/(2 r^5)) (quad_g5c.c:285-295 + phantom_g5.c:244-269).
 * The ...MC0 variants use a second, independent cell memory, as the reference's
 * Jcdata0 layout does (quad_g5c.c:19-196).
 */
#ifndef QUAD_UTIL
#define QUAD_UTIL

#if defined(__cplusplus)
extern "C" {
#endif

void g5c_set_nMC(int devid, int n);                                                          /* quad_g5c.c:14 */
void g5c_set_xmjMC0(int devid, int adr, int nj, double (*xj)[3], double *mj, double (*qj)[6]); /* quad_g5c.c:19 */
void g5c_calculate_force_on_xMC0(int devid, double (*x)[3], double (*a)[3], double *p, int ni); /* quad_g5c.c:178 */
void g5c_runMC0(int devid);                                                                  /* quad_g5c.c:190 */
void g5c_set_xmjMC(int devid, int adr, int nj, double (*xj)[3], double *mj, double (*qj)[6]);  /* quad_g5c.c:198 */
void g5c_calculate_force_on_xMC(int devid, double (*x)[3], double (*a)[3], double *p, int ni);  /* quad_g5c.c:285 */
void g5c_runMC(int devid);                                                                   /* quad_g5c.c:297 */

#if defined(__cplusplus)
}
#endif

#endif
