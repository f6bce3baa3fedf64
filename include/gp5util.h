/* GRAPE-5 API of libpg5.so (and libpg5_sym.so with -DSYMMETRIC).
 * pgamd-library: pg5 pg5_sym
 *
 * Same declarations as the reference header PG5/newton/libpg5/gp5util.h:1-45,
 * so existing callers compile unchanged.  Every function below is exported by
 * libpg5.so; the Fortran entry points (name + '_', all arguments by
 * reference) follow PG5/newton/libpg5/pg5_fortran.c:1-119.
 *
 * Semantics (phantom_g5.c:35-383): a = sum_j m_j (x_j - x_i) / (r^2+eps^2)^1.5
 * and p = + sum_j m_j / (r^2+eps^2)^0.5, the pair j = i included; callers
 * form phi = -p + m_i/eps (direct.c:165-172).  The rsqrt bias correction of
 * the reference (rsqrt.c:42-81) is 0 here: the GPU rsqrt is accurate to ~1 ulp.
 * JMEMSIZE is the reference's value; g5_get_jmemsize() returns this library's
 * larger j-memory limit.
 */
#ifndef GP5UTIL_H
#define GP5UTIL_H

#if defined(__cplusplus)
extern "C" {
#endif

#define JMEMSIZE (1<<16) /* gp5util.h:8 of the reference */

#ifdef SYMMETRIC
void g5_set_xmjMC(int devid, int adr, int nj, double (*xj)[3], double *mj, double *epsj); /* phantom_g5.c:141 */
void g5_set_xmj(int adr, int nj, double (*xj)[3], double *mj, double *epsj);              /* phantom_g5.c:210 */
#else
void g5_set_xmjMC(int devid, int adr, int nj, double (*xj)[3], double *mj); /* phantom_g5.c:188 */
void g5_set_xmj(int adr, int nj, double (*xj)[3], double *mj);              /* phantom_g5.c:215 */
#endif

void g5_set_nMC(int devid, int n);                                     /* phantom_g5.c:89 */
void g5_set_xiMC(int devid, int ni, double (*xi)[3]);                  /* phantom_g5.c:100 */
void g5_set_xeiMC(int devid, int ni, double (*xi)[3], double *eps);    /* phantom_g5.c:115 */
void g5_runMC(int devid);                                              /* phantom_g5.c:222 */
void g5_get_forceMC(int devid, int ni, double (*a)[3], double *p);     /* phantom_g5.c:244 */
void g5_calculate_force_on_xMC(int devid, double (*x)[3], double (*a)[3], double *p, int ni);  /* :276 */
void g5_calculate_force_on_xeMC(int devid, double (*x)[3], double *eps2, double (*a)[3], double *p, int ni); /* :289 */

void g5_open(void);                              /* phantom_g5.c:45 */
void g5_close(void);                             /* phantom_g5.c:62 */
int  g5_get_number_of_pipelines(void);           /* phantom_g5.c:35 */
int  g5_get_jmemsize(void);                      /* phantom_g5.c:40 */
void g5_set_n(int n);                            /* phantom_g5.c:95 */

void g5_set_eps_to_all(double eps);              /* phantom_g5.c:67 */
void g5_set_eps(int ni, double *eps);            /* phantom_g5.c:84 */
void g5_set_epsMC(int devid, int ni, double *eps); /* phantom_g5.c:72 (defined, undeclared there) */
void g5_set_xi(int ni, double (*xi)[3]);         /* phantom_g5.c:130 */
void g5_set_xei(int ni, double (*xi)[3], double *eps2); /* phantom_g5.c:135 */
void g5_run(void);                               /* phantom_g5.c:239 */
void g5_get_force(int ni, double (*a)[3], double *p);   /* phantom_g5.c:271 */
void g5_calculate_force_on_x(double (*x)[3], double (*a)[3], double *p, int ni);  /* :303,315 */
void g5_calculate_force_on_xe(double (*x)[3], double *eps2, double (*a)[3], double *p, int ni); /* :308,350 */

/* Fortran entry points, PG5/newton/libpg5/pg5_fortran.c */
int  g5_get_number_of_pipelines_(void);          /* pg5_fortran.c:5 */
int  g5_get_jmemsize_(void);                     /* pg5_fortran.c:10 */
void g5_open_(void);                             /* pg5_fortran.c:14 */
void g5_close_(void);                            /* pg5_fortran.c:19 */
void g5_set_eps_to_all_(double *eps);            /* pg5_fortran.c:23 */
void g5_set_n_(int *n);                          /* pg5_fortran.c:28 */
void g5_set_nMC_(int *devid, int *n);            /* pg5_fortran.c:33 */
void g5_set_xi_(int *ni, double (*xi)[3]);       /* pg5_fortran.c:38 */
void g5_set_xei_(int *ni, double (*xi)[3], double *eps2);            /* pg5_fortran.c:43 */
void g5_set_xiMC_(int *devid, int *ni, double (*xi)[3]);             /* pg5_fortran.c:48 */
void g5_set_xeiMC_(int *devid, int *ni, double (*xi)[3], double *eps2); /* pg5_fortran.c:53 */
#ifdef SYMMETRIC
void g5_set_xmj_(int *adr, int *nj, double (*xj)[3], double *mj, double *eps2);              /* :59 */
void g5_set_xmjMC_(int *devid, int *adr, int *nj, double (*xj)[3], double *mj, double *eps2); /* :71 */
#else
void g5_set_xmj_(int *adr, int *nj, double (*xj)[3], double *mj);                /* pg5_fortran.c:64 */
void g5_set_xmjMC_(int *devid, int *adr, int *nj, double (*xj)[3], double *mj);  /* pg5_fortran.c:76 */
#endif
void g5_run_(void);                              /* pg5_fortran.c:82 */
void g5_runMC_(int *devid);                      /* pg5_fortran.c:86 */
void g5_get_force_(int *ni, double (*a)[3], double *p);                /* pg5_fortran.c:91 */
void g5_get_forceMC_(int *devid, int *ni, double (*a)[3], double *p);  /* pg5_fortran.c:96 */
void g5_calculate_force_on_x_(double (*x)[3], double (*a)[3], double *p, int *ni);   /* :101 */
void g5_calculate_force_on_xe_(double (*x)[3], double *eps2, double (*a)[3], double *p, int *ni); /* :106 */
void g5_calculate_force_on_xMC_(int *devid, double (*x)[3], double (*a)[3], double *p, int *ni); /* :111 */
void g5_calculate_force_on_xeMC_(int *devid, double (*x)[3], double *eps2, double (*a)[3], double *p, int *ni); /* :116 */

#if defined(__cplusplus)
}
#endif

#endif /* GP5UTIL_H */
