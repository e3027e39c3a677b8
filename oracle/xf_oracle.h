/*
 * xf_oracle.h -- CPU oracle of the FDTD time-stepping path (test
 * infrastructure only; see xf_oracle.c).
 */
#ifndef XF_ORACLE_H_
#define XF_ORACLE_H_

#include "../include/xfdtd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Advance the scene n_steps in place (DomainHD::run order). 0 on success. */
int xfo_run(xf_scene* s, int64_t n_steps);
const char* xfo_last_error(void);
int xfo_num_threads(void);
void xfo_set_num_threads(int n);
/* One Yee component update: flat form, and written as the reference's
 * triple loop. */
int xfo_yee_component(xf_scene* s, int is_e, int xyz);
int xfo_yee_component_triple_loop(xf_scene* s, int is_e, int xyz);
/* The learned layer's column predictor alone (n columns of 12 inputs). */
int xfo_lb_predict(const uint16_t* weights, double mur_k, const float* x, int64_t n,
                   float* out);
/* The predictor's 16-term matrix-product model alone (n products). */
int xfo_lb_mfma16(const float* c, const float* a, const float* b, int64_t n, float* out);

#ifdef __cplusplus
}
#endif

#endif
