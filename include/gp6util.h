/* GRAPE-6 API of libpg6.so.
 * pgamd-library: pg6
 *
 * Same declarations as the reference header PG6/libpg6/gp6util.h:1-101.
 * Conventions (PG6/libpg6/phantom_g6.c, gravity.c):
 *   - g6_set_j_particle takes acc = a/2 and jrk = j/6 (PG6/banana/banana01.c:
 *     199-212) and predicts x + v dt + acc dt^2 + jrk dt^3 (gravity.c:19-20);
 *   - outputs: acc, jerk, pot = -sum m/(r^2+eps^2)^0.5 with the self pair
 *     excluded by particle index; nnb = index of the nearest j, rnnb = its
 *     distance sqrt(r^2+eps^2);
 *   - neighbour lists (lasthalf*n): indices with sqrt(r^2+eps^2) <= h, h^2 per
 *     i, at most 1024 per i (avx_type.h:2), in j-address order.
 * g6_npipes() is 48 as in the reference (avx_type.h:1), or PGAMD_G6_NPIPES;
 * lasthalf accepts any ni.  gpid is ignored (one board).
 */
#ifndef GP6UTIL_H
#define GP6UTIL_H

#if defined(__cplusplus)
extern "C" {
#endif

void g6_open(int);                               /* phantom_g6.c:18 */
void g6_open_(int *);                            /* phantom_g6.c:28 */
void g6_set_tunit(int);                          /* phantom_g6.c:36 */
void g6_set_tunit_(int *);                       /* phantom_g6.c:42 */
void g6_set_xunit(int);                          /* phantom_g6.c:50 */
void g6_set_xunit_(int *);                       /* phantom_g6.c:56 */
void g6_reset(int);                              /* phantom_g6.c:64 */
void g6_reset_(int *);                           /* phantom_g6.c:70 */
void g6_reset_fofpga(int);                       /* phantom_g6.c:78 */
void g6_reset_fofpga_(int *);                    /* phantom_g6.c:84 */
void g6_close(int);                              /* phantom_g6.c:92 */
void g6_close_(int *);                           /* phantom_g6.c:98 */
int g6_npipes(void);                             /* phantom_g6.c:106 */
int g6_npipes_(void);                            /* phantom_g6.c:111 */

void g6_set_j_particle(int, int, int, double, double, double,
                       double *, double *, double *, double *, double *);     /* phantom_g6.c:118 */
void g6_set_j_particle_(int *, int *, int *, double *, double *, double *,
                        double *, double *, double *, double *, double *);   /* phantom_g6.c:127 */

void g6_set_ti(int, double);                     /* phantom_g6.c:138 */
void g6_set_ti_(int *, double *);                /* phantom_g6.c:145 */

void g6calc_firsthalf(int, int, int, int *, double (*)[3], double (*)[3],
                      double (*)[3], double (*)[3], double *, double, double *);   /* phantom_g6.c:153 */
void g6calc_firsthalf_(int *, int *, int *, int *, double (*)[3], double (*)[3],
                       double (*)[3], double (*)[3], double *, double *, double *); /* phantom_g6.c:166 */

int g6calc_lasthalf(int, int, int, int *, double (*)[3], double (*)[3], double, double *,
                    double (*)[3], double (*)[3], double *);                       /* phantom_g6.c:177 */
int g6calc_lasthalf_(int *, int *, int *, int *, double (*)[3], double (*)[3], double *, double *,
                     double (*)[3], double (*)[3], double *);                      /* phantom_g6.c:213 */
int g6calc_lasthalf2(int, int, int, int *, double (*)[3], double (*)[3], double, double *,
                     double (*)[3], double (*)[3], double *, int *);               /* phantom_g6.c:223 */
int g6calc_lasthalf2_(int *, int *, int *, int *, double (*)[3], double (*)[3], double *, double *,
                      double (*)[3], double (*)[3], double *, int *);              /* phantom_g6.c:260 */
int g6calc_lasthalf2p(int, int, int, int *, double (*)[3], double (*)[3], double, double *,
                      double (*)[3], double (*)[3], double *, int *, float *);     /* phantom_g6.c:270 */
int g6calc_lasthalf2p_(int *, int *, int *, int *, double (*)[3], double (*)[3], double *, double *,
                       double (*)[3], double (*)[3], double *, int *, float *);    /* phantom_g6.c:308 */
int g6calc_lasthalfn(int, int, int, int *, double (*)[3], double (*)[3], double, double *,
                     double (*)[3], double (*)[3], double *);                      /* phantom_g6.c:318 */
int g6calc_lasthalfn_(int *, int *, int *, int *, double (*)[3], double (*)[3], double *, double *,
                      double (*)[3], double (*)[3], double *);                     /* phantom_g6.c:357 */
int g6calc_lasthalf2n(int, int, int, int *, double (*)[3], double (*)[3], double, double *,
                      double (*)[3], double (*)[3], double *, int *);              /* phantom_g6.c:367 */
int g6calc_lasthalf2n_(int *, int *, int *, int *, double (*)[3], double (*)[3], double *, double *,
                       double (*)[3], double (*)[3], double *, int *);             /* phantom_g6.c:407 */
int g6calc_lasthalf2np(int, int, int, int *, double (*)[3], double (*)[3], double, double *,
                       double (*)[3], double (*)[3], double *, int *, float *);    /* phantom_g6.c:417 */
int g6calc_lasthalf2np_(int *, int *, int *, int *, double (*)[3], double (*)[3], double *, double *,
                        double (*)[3], double (*)[3], double *, int *, float *);   /* phantom_g6.c:458 */

void g6_setup_njdata(int, int);                  /* phantom_g6.c:468 */
void g6_setup_njdata_(int *, int *);             /* phantom_g6.c:474 */
int g6_read_neighbour_list(int);                 /* phantom_g6.c:482 */
int g6_read_neighbour_list_(int *);              /* phantom_g6.c:488 */
int g6_get_neighbour_list(int, int, int, int *, int *);       /* phantom_g6.c:495 */
int g6_get_neighbour_list_(int *, int *, int *, int *, int *); /* phantom_g6.c:501 */
void g6_debugfunc(int);                          /* phantom_g6.c:508 */
void g6_debugfunc_(int *);                       /* phantom_g6.c:519 */
void g6_debugfunc_double(double);                /* phantom_g6.c:527 */
void g6_debugfunc_double_(double *);             /* phantom_g6.c:537 */
void g6_dump(double, int, double *, double (*)[3], double (*)[3]);      /* phantom_g6.c:545 */
void g6_dump_(double *, int *, double *, double (*)[3], double (*)[3]); /* phantom_g6.c:563 */

#if defined(__cplusplus)
}
#endif

#endif /* GP6UTIL_H */
