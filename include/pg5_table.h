/* Tabulated-cutoff GRAPE-5 API of libpg5_table.so (TreePM short range).
 * pgamd-library: pg5_table
 *
 * Declarations of PG5/table/libpg5/gp5util.h:52-92 and pg5_table.h:31-43.
 * This library exports the same g5_* names as libpg5.so with the table
 * semantics, so the two are separate shared objects, as in the reference.
 *
 * Semantics (phantom_g5.c:43-200): positions are scaled by xscale =
 * sqrt(r2scale); a = sum_j m_j f(r) (x_j - x_i) with f read from a
 * 2^(4+FRC_BIT)-entry {value, slope} table indexed by the float bits of
 * s = 2 + r^2 r2scale (gravity_kernel_avx2.c:22-137); p = 0.  Default shape:
 * the S2 short-range force with eps_PP = 0.1/64, eps_PM = r_cut = 3/64
 * (pppm.h:4-6), built by g5_open.
 */
#ifndef PG5_TABLE_API_H
#define PG5_TABLE_API_H

#if defined(__cplusplus)
extern "C" {
#endif

/* pg5_table.h:31 — the table of the default 512-entry resolution */
extern float Force_table[512][2];

void pg5_gen_force_table(double (*force_func)(double), double rcut); /* pg5_table.c:43 */
void pg5_gen_s2_force_table(double sft_for_PP, double sft_for_PM);   /* pg5_table.c:31 */
void pg5_set_xscale(double xscale);                                  /* phantom_g5.c:43 */

void g5_set_xmjMC(int devid, int adr, int nj, double (*xj)[3], double *mj); /* phantom_g5.c:112 */
void g5_set_nMC(int devid, int n);                                          /* phantom_g5.c:86 */
void g5_set_xiMC(int devid, int ni, double (*xi)[3]);                       /* phantom_g5.c:96 */
void g5_runMC(int devid);                                                   /* phantom_g5.c:134 */
void g5_get_forceMC(int devid, int ni, double (*a)[3], double *p);          /* phantom_g5.c:145 */

void g5_open(void);                                                 /* phantom_g5.c:56 */
void g5_close(void);                                                /* phantom_g5.c:63 */
int  g5_get_number_of_pipelines(void);                              /* phantom_g5.c:52 */
int  g5_get_jmemsize(void);                                         /* (newton gp5util.h:29) */
void g5_set_range(double xmin, double xmax, double mmin);           /* phantom_g5.c:77 */
void g5_set_n(int n);                                               /* phantom_g5.c:92 */
void g5_set_mj(int adr, int nj, double *mj);                        /* gp5util.h:72 (declared only there) */
void g5_set_xj(int adr, int nj, double (*xj)[3]);                   /* gp5util.h:73 (declared only there) */
void g5_set_xmj(int adr, int nj, double (*xj)[3], double *mj);      /* phantom_g5.c:130 */
void g5_set_eps_to_all(double eps);                                 /* phantom_g5.c:73 */
void g5_set_xi(int ni, double (*xi)[3]);                            /* phantom_g5.c:108 */
void g5_run(void);                                                  /* phantom_g5.c:141 */
void g5_get_force(int ni, double (*a)[3], double *p);               /* phantom_g5.c:156 */
void g5_set_cutoff_table(double (*ffunc)(double), double fcut, double fcor,
                         double (*pfunc)(double), double pcut, double pcor); /* phantom_g5.c:80 */
void g5_set_eta(double eta);                                        /* phantom_g5.c:69 */
void g5_calculate_force_on_xMC(int devid, double (*x)[3], double (*a)[3], double *p, int ni); /* :160 */
void g5_calculate_force_on_x(double (*x)[3], double (*a)[3], double *p, int ni);              /* :173,179 */

#if defined(__cplusplus)
}
#endif

#endif
