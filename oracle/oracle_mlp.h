/* ORACLE - test infrastructure, not product code.
 * CPU restatement of the learned con2prim (oracle_mlp.c). */
#ifndef ECHO_ORACLE_MLP_H
#define ECHO_ORACLE_MLP_H
#ifdef __cplusplus
extern "C" {
#endif
#define MLP_NIN 8
#define MLP_NH 64
#define MLP_NOUT 5
#define MLP_NPARAM 9221

double orc_f16(double x);
double orc_act_h(float z);
float orc_act(float x);
float orc_act_d(float x);
void orc_mlp_features(const double *u, double g33, float *x);
double orc_mlp_w0(const double *u, double g33);
void orc_mlp_targets(const double *u, const double *v, double g33, double *t);
void orc_mlp_forward(const float *P, const float *x, int round, float *y);
double orc_mlp_loss(const float *P, const double *u, const double *v, long n, double g33,
                    int round, double *grad);
void orc_e3p_w(double W, double S2, double B2, double SB2, double U, double *out3);
int orc_mlp_refine(const double *u_in, const float *y, double g33, double gm, int iters,
                   double tol, double *v);
#ifdef __cplusplus
}
#endif
#endif
