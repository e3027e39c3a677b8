/*
 * xf_oracle.c -- CPU restatement of XFDTD_CUDA's per-timestep FDTD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * xfdtd_cuda_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never calls it.
 *
 * It advances an xf_scene (include/xfdtd_hip.h) in place, following the
 * reference kernels operation by operation, in the reference's step order
 * (src/domain/domain_hd.cu:35-87):
 *   updateH -> correctH (TFSF, then PML X, Y, Z) -> updateE ->
 *   correctE (TFSF, then PML) -> record (movies, NF2FF-FD, NF2FF-TD) ->
 *   nextStep.
 * Floating-point expressions keep the reference's evaluation order and are
 * compiled without contraction (-ffp-contract=off), so a GPU build with the
 * same order and no FMA reproduces the result bit for bit.
 *
 * Pins (see tests/): Yee coefficients and dispersion against closed forms,
 * PEC-cavity energy, CPML reflection, TFSF leakage, the DFT of a known
 * sinusoid for NF2FF, the telescoping sums and first moments of the
 * NF2FF-TD potentials, the analytic permittivity of the ADE media, and
 * a brute-force triple-loop form of the flat Yee update.
 */
#include "xf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

/* ---------------------------------------------------------------- helpers */

static inline double* at3(const xf_array* a, int64_t i, int64_t j, int64_t k) {
  return a->data + ((i * a->shape[1] + j) * a->shape[2] + k);
}

static inline double* at4(const xf_array* a, int64_t i, int64_t j, int64_t k,
                          int64_t p) {
  return a->data + (((i * a->shape[1] + j) * a->shape[2] + k) * a->shape[3] + p);
}

/* complex arrays: trailing dimension 2 */
static inline double* atc3(const xf_array* a, int64_t i, int64_t j, int64_t k) {
  return a->data + (((i * a->shape[1] + j) * a->shape[2] + k) * 2);
}

static inline double at2(const xf_array* a, int64_t r, int64_t c) {
  return a->data[r * a->shape[1] + c];
}

/* transform::xYZToABC / aBCToXYZ (XFDTD_CORE util/transform/abc_xyz.h, used
 * at update_scheme.cuh:172-178): X -> (a,b,c) = (y,z,x), Y -> (z,x,y),
 * Z -> (x,y,z). */
static inline void xyz_to_abc(int axis, int64_t i, int64_t j, int64_t k,
                              int64_t* a, int64_t* b, int64_t* c) {
  if (axis == 0) { *a = j; *b = k; *c = i; }
  else if (axis == 1) { *a = k; *b = i; *c = j; }
  else { *a = i; *b = j; *c = k; }
}

static inline void abc_to_xyz(int axis, int64_t a, int64_t b, int64_t c,
                              int64_t* i, int64_t* j, int64_t* k) {
  if (axis == 0) { *i = c; *j = a; *k = b; }
  else if (axis == 1) { *i = b; *j = c; *k = a; }
  else { *i = a; *j = b; *k = c; }
}

static inline int64_t range_get_start(const xf_task* t, int axis) {
  return axis == 0 ? t->x.start : axis == 1 ? t->y.start : t->z.start;
}
static inline int64_t range_get_end(const xf_task* t, int axis) {
  return axis == 0 ? t->x.end : axis == 1 ? t->y.end : t->z.end;
}

/* Coefficient n (order of xf_scene.coeff) at (i, j, k): the dense array, or
 * the compact material table with clamped cell indices. */
static inline double coef(const xf_scene* s, int n, int64_t i, int64_t j,
                          int64_t k) {
  if (s->coeff[n].data) return *at3(&s->coeff[n], i, j, k);
  const int64_t ci = i < s->nx ? i : s->nx - 1;
  const int64_t cj = j < s->ny ? j : s->ny - 1;
  const int64_t ck = k < s->nz ? k : s->nz - 1;
  return s->material_coeff[(int64_t)s->cell_material[(ci * s->ny + cj) * s->nz + ck] * 18 + n];
}

static const xf_array* efield(const xf_scene* s, int axis) {
  return axis == 0 ? &s->ex : axis == 1 ? &s->ey : &s->ez;
}
static const xf_array* hfield(const xf_scene* s, int axis) {
  return axis == 0 ? &s->hx : axis == 1 ? &s->hy : &s->hz;
}

/* ---------------------------------------------------------- Yee updates */

/* eNext / hNext: update_scheme.cuh:27-32, 58-63. */
static inline double e_next(double cece, double ec, double cecha, double ha_p,
                            double ha_q, double cechb, double hb_p,
                            double hb_q) {
  return cece * ec + cecha * (ha_p - ha_q) + cechb * (hb_p - hb_q);
}
static inline double h_next(double chch, double hc, double chcea, double ea_p,
                            double ea_q, double chceb, double eb_p,
                            double eb_q) {
  return chch * hc + chcea * (ea_q - ea_p) + chceb * (eb_q - eb_p);
}

/* update<attribute, xyz>(emf, coeff, start, end, nx, ny, nz):
 * update_scheme.cuh:141-197 over the whole node range [0, nx*ny*nz). */
static void yee_update(const xf_scene* s, int is_e, int xyz) {
  const int a_axis = (xyz + 1) % 3, b_axis = (xyz + 2) % 3;
  const xf_array* f = is_e ? efield(s, xyz) : hfield(s, xyz);
  const xf_array* fa = is_e ? hfield(s, a_axis) : efield(s, a_axis);
  const xf_array* fb = is_e ? hfield(s, b_axis) : efield(s, b_axis);
  const int base = (is_e ? 0 : 9) + 3 * xyz;
  const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
  const int64_t off = is_e ? -1 : 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nx; ++i) {
    for (int64_t j = 0; j < ny; ++j) {
      for (int64_t k = 0; k < nz; ++k) {
        int64_t a, b, c, ia, ja, ka, ib, jb, kb;
        xyz_to_abc(xyz, i, j, k, &a, &b, &c);
        abc_to_xyz(xyz, a, b + off, c, &ia, &ja, &ka);
        abc_to_xyz(xyz, a + off, b, c, &ib, &jb, &kb);
        if (is_e) {
          if (a == 0 || b == 0) continue;
          *at3(f, i, j, k) = e_next(coef(s, base, i, j, k), *at3(f, i, j, k),
                                    coef(s, base + 1, i, j, k), *at3(fa, i, j, k),
                                    *at3(fa, ia, ja, ka), coef(s, base + 2, i, j, k),
                                    *at3(fb, i, j, k), *at3(fb, ib, jb, kb));
        } else {
          *at3(f, i, j, k) = h_next(coef(s, base, i, j, k), *at3(f, i, j, k),
                                    coef(s, base + 1, i, j, k), *at3(fa, i, j, k),
                                    *at3(fa, ia, ja, ka), coef(s, base + 2, i, j, k),
                                    *at3(fb, i, j, k), *at3(fb, ib, jb, kb));
        }
      }
    }
  }
}

/* ADE E update: TemplateADEUpdateScheme::updateE
 * (template_ade_update_scheme.cuh:15-86) with the model's calculateJSum /
 * updateJ / ePrevious / coeffEPrev / recordEPrevious (drude_ade_updator.cu:
 * 35-60, debye_ade_updator.cu:27-52, m_lor_ade_updator.cu:27-88, and the
 * zero-returning ePrevious/coeffEPrev of drude/debye_ade_updator.cuh). */
static void ade_update_e(const xf_scene* s, int xyz) {
  const xf_ade* ad = s->ade;
  const int a_axis = (xyz + 1) % 3, b_axis = (xyz + 2) % 3;
  const xf_array* f = efield(s, xyz);
  const xf_array* fa = hfield(s, a_axis);
  const xf_array* fb = hfield(s, b_axis);
  const int base = 3 * xyz;
  const xf_array* jarr = xyz == 0 ? &ad->jx : xyz == 1 ? &ad->jy : &ad->jz;
  const xf_array* jparr =
      xyz == 0 ? &ad->jx_prev : xyz == 1 ? &ad->jy_prev : &ad->jz_prev;
  const xf_array* eprev =
      xyz == 0 ? &ad->ex_prev : xyz == 1 ? &ad->ey_prev : &ad->ez_prev;
  const int64_t np = ad->num_pole;
  const int kind = ad->kind;
  const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nx; ++i) {
    for (int64_t j = 0; j < ny; ++j) {
      for (int64_t k = 0; k < nz; ++k) {
        int64_t a, b, c, ia, ja, ka, ib, jb, kb;
        xyz_to_abc(xyz, i, j, k, &a, &b, &c);
        abc_to_xyz(xyz, a, b - 1, c, &ia, &ja, &ka);
        abc_to_xyz(xyz, a - 1, b, c, &ib, &jb, &kb);
        if (a == 0 || b == 0) continue;
        /* calculateJSum */
        double j_sum = 0;
        for (int64_t p = 0; p < np; ++p) {
          if (kind == XF_ADE_MLORENTZ) {
            j_sum += *at4(&ad->coeff_j_sum_j, i, j, k, p) * *at4(jarr, i, j, k, p) +
                     *at4(&ad->coeff_j_sum_j_p, i, j, k, p) * *at4(jparr, i, j, k, p);
          } else {
            j_sum += *at4(&ad->coeff_j_sum_j, i, j, k, p) * *at4(jarr, i, j, k, p);
          }
        }
        double e_prev = 0.0, coeff_e_e_p = 0.0;
        if (kind == XF_ADE_MLORENTZ) {
          e_prev = *at3(eprev, i, j, k);
          coeff_e_e_p = *at3(&ad->coeff_e_e_p, i, j, k);
        }
        const double e_cur = *at3(f, i, j, k);
        const double e_new =
            coeff_e_e_p * e_prev +
            e_next(coef(s, base, i, j, k), *at3(f, i, j, k), coef(s, base + 1, i, j, k),
                   *at3(fa, i, j, k), *at3(fa, ia, ja, ka), coef(s, base + 2, i, j, k),
                   *at3(fb, i, j, k), *at3(fb, ib, jb, kb)) +
            *at3(&ad->coeff_e_j_sum, i, j, k) * j_sum;
        *at3(f, i, j, k) = e_new;
        /* updateJ */
        for (int64_t p = 0; p < np; ++p) {
          double* jv = at4(jarr, i, j, k, p);
          if (kind == XF_ADE_DRUDE) {
            *jv = *at4(&ad->coeff_j_j, i, j, k, p) * *jv +
                  *at4(&ad->coeff_j_e, i, j, k, p) * (e_new + e_cur);
          } else if (kind == XF_ADE_DEBYE) {
            *jv = *at4(&ad->coeff_j_j, i, j, k, p) * *jv +
                  *at4(&ad->coeff_j_e, i, j, k, p) * (e_new - e_cur);
          } else {
            double* jpv = at4(jparr, i, j, k, p);
            const double j_next =
                *at4(&ad->coeff_j_e_n, i, j, k, p) * e_new +
                *at4(&ad->coeff_j_e, i, j, k, p) * e_cur +
                *at4(&ad->coeff_j_e_p, i, j, k, p) * e_prev +
                *at4(&ad->coeff_j_j, i, j, k, p) * *jv +
                *at4(&ad->coeff_j_j_p, i, j, k, p) * *jpv;
            *jpv = *jv;
            *jv = j_next;
          }
        }
        /* recordEPrevious */
        if (kind == XF_ADE_MLORENTZ) *at3(eprev, i, j, k) = e_cur;
      }
    }
  }
}

/* ------------------------------------------------------------------ CPML */

/* correctPML<attribute, xyz> (pml_corrector.cu:9-48) over the box
 * a in [a0,a1), b in [b0,b1), c in [c0,c1). */
static void correct_pml_box(int is_e, int xyz, const xf_array* field,
                            const xf_array* psi, const xf_array* coeff_a,
                            const xf_array* coeff_b, const xf_array* dual,
                            const xf_array* c_psi, int64_t a0, int64_t a1,
                            int64_t b0, int64_t b1, int64_t c0, int64_t c1,
                            int64_t pml_global_start, int64_t pml_node_start,
                            int64_t offset_c) {
  const int64_t off = is_e ? -1 : 1;
#pragma omp parallel for schedule(static) collapse(2)
  for (int64_t c = c0; c < c1; ++c) {
    for (int64_t a = a0; a < a1; ++a) {
      for (int64_t b = b0; b < b1; ++b) {
        int64_t i, j, k, il, jl, kl, id, jd, kd;
        abc_to_xyz(xyz, a, b, c, &i, &j, &k);
        const int64_t global_c = c + offset_c;
        const double ca = coeff_a->data[global_c - pml_global_start];
        const double cb = coeff_b->data[global_c - pml_global_start];
        const int64_t layer = c - pml_node_start;
        abc_to_xyz(xyz, a, b, layer, &il, &jl, &kl);
        const double cpsi = *at3(c_psi, il, jl, kl);
        double* pv = at3(psi, il, jl, kl);
        double* fv = at3(field, i, j, k);
        const double dp = *at3(dual, i, j, k);
        abc_to_xyz(xyz, a, b, c + off, &id, &jd, &kd);
        const double dq = *at3(dual, id, jd, kd);
        if (is_e) {
          *pv = cb * *pv + ca * (dp - dq);
        } else {
          *pv = cb * *pv + ca * (dq - dp);
        }
        *fv += cpsi * *pv;
      }
    }
  }
}

/* PMLCorrector<xyz>::correctE / correctH (pml_corrector.cu:50-270): the union
 * of all thread sub-tasks; the "last block" lateral +1 extension becomes an
 * extension of the whole range. */
static void pml_correct(const xf_scene* s, const xf_pml* p, int is_e) {
  const int xyz = p->main_axis;
  const int a_axis = (xyz + 1) % 3, b_axis = (xyz + 2) % 3;
  int64_t as, bs, cs, ae, be, ce;
  xyz_to_abc(xyz, p->task.x.start, p->task.y.start, p->task.z.start, &as, &bs,
             &cs);
  xyz_to_abc(xyz, p->task.x.end, p->task.y.end, p->task.z.end, &ae, &be, &ce);
  if (!(p->task.x.start < p->task.x.end && p->task.y.start < p->task.y.end &&
        p->task.z.start < p->task.z.end))
    return;
  if (is_e) {
    const int64_t off = (cs == 0) ? 1 : 0;
    /* EA: [a_s, a_e), [b_s, b_e + 1) */
    correct_pml_box(1, xyz, efield(s, a_axis), &p->ea_psi_hb, &p->coeff_a_e,
                    &p->coeff_b_e, hfield(s, b_axis), &p->c_ea_psi_hb, as, ae,
                    bs, be + 1, cs + off, ce + off, p->global_e_start,
                    p->node_e_start, p->offset_c);
    /* EB: [a_s, a_e + 1), [b_s, b_e) */
    correct_pml_box(1, xyz, efield(s, b_axis), &p->eb_psi_ha, &p->coeff_a_e,
                    &p->coeff_b_e, hfield(s, a_axis), &p->c_eb_psi_ha, as,
                    ae + 1, bs, be, cs + off, ce + off, p->global_e_start,
                    p->node_e_start, p->offset_c);
  } else {
    /* HA: [a_s, a_e + 1), [b_s, b_e) */
    correct_pml_box(0, xyz, hfield(s, a_axis), &p->ha_psi_eb, &p->coeff_a_h,
                    &p->coeff_b_h, efield(s, b_axis), &p->c_ha_psi_eb, as,
                    ae + 1, bs, be, cs, ce, p->global_h_start, p->node_h_start,
                    p->offset_c);
    /* HB: [a_s, a_e), [b_s, b_e + 1) */
    correct_pml_box(0, xyz, hfield(s, b_axis), &p->hb_psi_ea, &p->coeff_a_h,
                    &p->coeff_b_h, efield(s, a_axis), &p->c_hb_psi_ea, as, ae,
                    bs, be + 1, cs, ce, p->global_h_start, p->node_h_start,
                    p->offset_c);
  }
}

/* ------------------------------------------------------------------ TFSF */

/* exInc..hzInc (tfsf_corrector.cu:18-130). */
static double tfsf_inc(const xf_tfsf* tf, int is_e, int comp, int64_t t,
                       int64_t i, int64_t j, int64_t k) {
  const int64_t gi = tf->total_task.x.start, gj = tf->total_task.y.start,
                gk = tf->total_task.z.start;
  double projection;
  if (is_e) {
    if (comp == 0) {
      projection = tf->proj_x_half.data[i - gi + 1] +
                   tf->proj_y_int.data[j - gj] + tf->proj_z_int.data[k - gk];
    } else if (comp == 1) {
      projection = tf->proj_x_int.data[i - gi] +
                   tf->proj_y_half.data[j - gj + 1] +
                   tf->proj_z_int.data[k - gk];
    } else {
      projection = tf->proj_x_int.data[i - gi] + tf->proj_y_int.data[j - gj] +
                   tf->proj_z_half.data[k - gk + 1];
    }
  } else {
    if (comp == 0) {
      projection = tf->proj_x_int.data[i - gi] +
                   tf->proj_y_half.data[j - gj + 1] +
                   tf->proj_z_half.data[k - gk + 1] - 0.5;
    } else if (comp == 1) {
      projection = tf->proj_x_half.data[i - gi + 1] +
                   tf->proj_y_int.data[j - gj] +
                   tf->proj_z_half.data[k - gk + 1] - 0.5;
    } else {
      projection = tf->proj_x_half.data[i - gi + 1] +
                   tf->proj_y_half.data[j - gj + 1] +
                   tf->proj_z_int.data[k - gk] - 0.5;
    }
  }
  const size_t index = (size_t)projection;
  const double weight = projection - (double)index;
  const xf_array* tab = is_e ? &tf->e_inc : &tf->h_inc;
  const double tr = is_e ? tf->transform_e[comp] : tf->transform_h[comp];
  return tr * ((1 - weight) * at2(tab, t, (int64_t)index) +
               weight * at2(tab, t, (int64_t)index + 1));
}

/* correctTFSF<direction, attribute, field_xyz> (tfsf_corrector.cuh:20-170)
 * over the union of the thread tasks of one face. face: 0..5 = XN..ZP. */
static void tfsf_correct_face(const xf_scene* s, int face, int is_e,
                              int field_xyz, int64_t t) {
  const xf_tfsf* tf = s->tfsf;
  const int xyz = face / 2;
  const int negative = (face % 2) == 0;
  const int a_axis = (xyz + 1) % 3, b_axis = (xyz + 2) % 3;
  const int dual_field_xyz = (field_xyz == a_axis) ? b_axis : a_axis;
  const xf_task* tt = &tf->total_task;
  /* buildTask<direction> */
  xf_task task = *tt;
  const int64_t cpos = negative ? range_get_start(tt, xyz) : range_get_end(tt, xyz);
  if (xyz == 0) { task.x.start = cpos; task.x.end = cpos + 1; }
  else if (xyz == 1) { task.y.start = cpos; task.y.end = cpos + 1; }
  else { task.z.start = cpos; task.z.end = cpos + 1; }
  if (!(task.x.start < task.x.end && task.y.start < task.y.end &&
        task.z.start < task.z.end))
    return;
  int64_t as, bs, cs, ae, be, ce;
  xyz_to_abc(xyz, task.x.start, task.y.start, task.z.start, &as, &bs, &cs);
  xyz_to_abc(xyz, task.x.end, task.y.end, task.z.end, &ae, &be, &ce);
  const int offset_a = ((is_e && field_xyz == b_axis) || (!is_e && field_xyz == a_axis)) ? 1 : 0;
  const int offset_b = offset_a == 1 ? 0 : 1;
  if (offset_a == 1 && ae == range_get_end(tt, a_axis)) ae += offset_a;
  if (offset_b == 1 && be == range_get_end(tt, b_axis)) be += offset_b;
  const int equation_flag = is_e ? 1 : -1;
  const int different_flag = (field_xyz == a_axis) ? -1 : 1;
  const int direction_flag = negative ? -1 : 1;
  const int coefficient_flag = equation_flag * different_flag * direction_flag;
  const double cax[3] = {tf->cax, tf->cay, tf->caz};
  const double cbx[3] = {tf->cbx, tf->cby, tf->cbz};
  const double coefficient = (is_e ? cax[xyz] : cbx[xyz]) * coefficient_flag;
  const xf_array* field = is_e ? efield(s, field_xyz) : hfield(s, field_xyz);
  int64_t is, js, ks, ie, je, ke;
  abc_to_xyz(xyz, as, bs, cs, &is, &js, &ks);
  abc_to_xyz(xyz, ae, be, ce, &ie, &je, &ke);
  for (int64_t i = is; i < ie; ++i) {
    for (int64_t j = js; j < je; ++j) {
      for (int64_t k = ks; k < ke; ++k) {
        int64_t a, b, c;
        xyz_to_abc(xyz, i, j, k, &a, &b, &c);
        if (negative) {
          if (is_e) {
            int64_t id, jd, kd;
            abc_to_xyz(xyz, a, b, c - 1, &id, &jd, &kd);
            const double v = tfsf_inc(tf, 0, dual_field_xyz, t, id, jd, kd);
            *at3(field, i, j, k) += v * coefficient;
          } else {
            const double v = tfsf_inc(tf, 1, dual_field_xyz, t, i, j, k);
            int64_t ii, jj, kk;
            abc_to_xyz(xyz, a, b, c - 1, &ii, &jj, &kk);
            *at3(field, ii, jj, kk) += v * coefficient;
          }
          continue;
        }
        const double v = tfsf_inc(tf, !is_e, dual_field_xyz, t, i, j, k);
        *at3(field, i, j, k) += v * coefficient;
      }
    }
  }
}

/* TFSFCorrector3DAgency / TFSFCorrector2DAgency::correctE/H launch order
 * (tfsf_corrector.cu:155-271). */
static void tfsf_correct(const xf_scene* s, int is_e, int64_t t) {
  if (s->tfsf->dimension == 2) {
    if (is_e) {
      tfsf_correct_face(s, XF_XN, 1, 2, t);
      tfsf_correct_face(s, XF_XP, 1, 2, t);
      tfsf_correct_face(s, XF_YN, 1, 2, t);
      tfsf_correct_face(s, XF_YP, 1, 2, t);
    } else {
      tfsf_correct_face(s, XF_XN, 0, 1, t);
      tfsf_correct_face(s, XF_XP, 0, 1, t);
      tfsf_correct_face(s, XF_YN, 0, 0, t);
      tfsf_correct_face(s, XF_YP, 0, 0, t);
    }
    return;
  }
  const int e = is_e;
  tfsf_correct_face(s, XF_XN, e, 2, t);
  tfsf_correct_face(s, XF_XP, e, 2, t);
  tfsf_correct_face(s, XF_YN, e, 2, t);
  tfsf_correct_face(s, XF_YP, e, 2, t);
  tfsf_correct_face(s, XF_ZN, e, 0, t);
  tfsf_correct_face(s, XF_ZP, e, 0, t);
  tfsf_correct_face(s, XF_YN, e, 0, t);
  tfsf_correct_face(s, XF_YP, e, 0, t);
  tfsf_correct_face(s, XF_ZN, e, 1, t);
  tfsf_correct_face(s, XF_ZP, e, 1, t);
  tfsf_correct_face(s, XF_XN, e, 1, t);
  tfsf_correct_face(s, XF_XP, e, 1, t);
}

/* ----------------------------------------------------------------- NF2FF */

/* interpolateSurfaceCenter<xyz, field> (interpolate_scheme.cuh:11-134). */
static double interp(const xf_scene* s, int face_axis, int is_e, int comp,
                     int64_t i, int64_t j, int64_t k) {
  const xf_array* f = is_e ? efield(s, comp) : hfield(s, comp);
#define F(ii, jj, kk) (*at3(f, (ii), (jj), (kk)))
  if (face_axis == 0) {
    if (!is_e && comp == 2)
      return 0.25 * (F(i, j, k + 1) + F(i, j, k) + F(i - 1, j, k + 1) + F(i - 1, j, k));
    if (!is_e && comp == 1)
      return 0.25 * (F(i, j + 1, k) + F(i, j, k) + F(i - 1, j + 1, k) + F(i - 1, j, k));
    if (is_e && comp == 2) return 0.5 * (F(i, j + 1, k) + F(i, j, k));
    if (is_e && comp == 1) return 0.5 * (F(i, j, k + 1) + F(i, j, k));
  } else if (face_axis == 1) {
    if (!is_e && comp == 2)
      return 0.25 * (F(i, j, k + 1) + F(i, j, k) + F(i, j - 1, k + 1) + +F(i, j - 1, k));
    if (!is_e && comp == 0)
      return 0.25 * (F(i + 1, j, k) + F(i, j, k) + F(i + 1, j - 1, k) + F(i, j - 1, k));
    if (is_e && comp == 2) return 0.5 * (F(i + 1, j, k) + F(i, j, k));
    if (is_e && comp == 0) return 0.5 * (F(i, j, k + 1) + F(i, j, k));
  } else {
    if (!is_e && comp == 0)
      return 0.25 * (F(i + 1, j, k) + F(i, j, k) + F(i, j, k - 1) + F(i + 1, j, k - 1));
    if (!is_e && comp == 1)
      return 0.25 * (F(i, j, k) + F(i, j + 1, k) + F(i, j, k - 1) + F(i, j + 1, k - 1));
    if (is_e && comp == 0) return 0.5 * (F(i, j + 1, k) + F(i, j, k));
    if (is_e && comp == 1) return 0.5 * (F(i + 1, j, k) + F(i, j, k));
  }
#undef F
  return 0.0;
}

/* NF2FFFrequencyDomainData<D>::calculateJ / calculateM
 * (nf2ff_frequency_domain.data.cu:80-194). */
static void nf2ff_face(const xf_scene* s, const xf_nf2ff* nf,
                       const xf_nf2ff_freq* fr, int face, int64_t t) {
  const xf_task* task = &nf->face_task[face];
  if (!(task->x.start < task->x.end && task->y.start < task->y.end &&
        task->z.start < task->z.end))
    return;
  const int xyz = face / 2;
  const int positive = face % 2;
  const int a_axis = (xyz + 1) % 3, b_axis = (xyz + 2) % 3;
  const double coff_ja = positive ? -1.0 : 1.0;
  const double coff_jb = positive ? 1.0 : -1.0;
  const double coff_ma = positive ? 1.0 : -1.0;
  const double coff_mb = positive ? -1.0 : 1.0;
  const double* th = fr->transform_h.data + 2 * t;
  const double* te = fr->transform_e.data + 2 * t;
  const xf_array* ja = &fr->cur[face][0];
  const xf_array* jb = &fr->cur[face][1];
  const xf_array* ma = &fr->cur[face][2];
  const xf_array* mb = &fr->cur[face][3];
  const int64_t oi = task->x.start, oj = task->y.start, ok = task->z.start;
  for (int64_t i = task->x.start; i < task->x.end; ++i) {
    for (int64_t j = task->y.start; j < task->y.end; ++j) {
      for (int64_t k = task->z.start; k < task->z.end; ++k) {
        double r;
        double* p;
        r = coff_ja * interp(s, xyz, 0, b_axis, i, j, k);
        p = atc3(ja, i - oi, j - oj, k - ok);
        p[0] += r * th[0];
        p[1] += r * th[1];
        r = coff_jb * interp(s, xyz, 0, a_axis, i, j, k);
        p = atc3(jb, i - oi, j - oj, k - ok);
        p[0] += r * th[0];
        p[1] += r * th[1];
      }
    }
  }
  for (int64_t i = task->x.start; i < task->x.end; ++i) {
    for (int64_t j = task->y.start; j < task->y.end; ++j) {
      for (int64_t k = task->z.start; k < task->z.end; ++k) {
        double r;
        double* p;
        r = coff_ma * interp(s, xyz, 1, b_axis, i, j, k);
        p = atc3(ma, i - oi, j - oj, k - ok);
        p[0] += r * te[0];
        p[1] += r * te[1];
        r = coff_mb * interp(s, xyz, 1, a_axis, i, j, k);
        p = atc3(mb, i - oi, j - oj, k - ok);
        p[0] += r * te[0];
        p[1] += r * te[1];
      }
    }
  }
}

/* NF2FFTimeDomainData<D>::update<E|H> (nf2ff_time_domain_data.cu:103-188)
 * for one face, serially in (i, j, k) order as the reference's one-thread
 * launch (:213-223) runs it.  r is rVector (:54-68), ds surfaceArea
 * (:70-80), the prev arrays are indexed (a, b) (:81-186).  Returns nonzero
 * when a time bin falls outside the potential arrays. */
static int nf2ff_td_face(const xf_scene* s, xf_nf2ff_td* td, int face, int is_e, int64_t t) {
  const xf_task* task = &td->face_task[face];
  if (!(task->x.start < task->x.end && task->y.start < task->y.end &&
        task->z.start < task->z.end))
    return 0;
  const int c = face / 2;
  const int positive = face % 2;
  const int a = (c + 1) % 3, b = (c + 2) % 3;
  const double offset = is_e ? 0.5 : 0.0;
  double coeff_a = is_e ? 1.0 : -1.0;
  double coeff_b = is_e ? -1.0 : 1.0;
  if (!positive) {
    coeff_a *= -1.0;
    coeff_b *= -1.0;
  }
  const double dmax = is_e ? td->dmax_e : td->dmax_h;
  const double c_dt = 299792458.0 * s->dt;
  xf_array* pa = &td->pot[face][is_e ? 2 : 0];
  xf_array* pb = &td->pot[face][is_e ? 3 : 1];
  xf_array* qa = &td->prev[face][is_e ? 0 : 2];
  xf_array* qb = &td->prev[face][is_e ? 1 : 3];
  const int64_t n_pot = pa->shape[0];
  const int64_t o[3] = {task->x.start, task->y.start, task->z.start};
  const int64_t nb = qa->shape[1];
  for (int64_t i = task->x.start; i < task->x.end; ++i) {
    for (int64_t j = task->y.start; j < task->y.end; ++j) {
      for (int64_t k = task->z.start; k < task->z.end; ++k) {
        const int64_t idx[3] = {i, j, k};
        double r[3];
        for (int ax = 0; ax < 3; ++ax)
          r[ax] = ax == c ? td->e_node[ax].data[idx[ax]] : td->h_node[ax].data[idx[ax]];
        const double time_delay =
            dmax - (r[0] * td->r_unit[0] + r[1] * td->r_unit[1] + r[2] * td->r_unit[2]);
        const double time_step_delay = time_delay / c_dt;
        const double sn = time_step_delay + offset + (double)t;
        const int64_t nn = (int64_t)floor(sn);
        const double coeff_f = sn - (double)nn;
        if (nn < 0 || nn + 1 >= n_pot) return 1;
        const double ds = td->e_size[a].data[idx[a]] * td->e_size[b].data[idx[b]];
        const double a_avg = interp(s, c, is_e, a, i, j, k);
        const double b_avg = interp(s, c, is_e, b, i, j, k);
        const int64_t lin = (idx[a] - o[a]) * nb + (idx[b] - o[b]);
        const double delta_a = a_avg - qa->data[lin];
        const double delta_b = b_avg - qb->data[lin];
        pa->data[nn] += coeff_a * (1 - coeff_f) * ds * delta_b;
        pa->data[nn + 1] += coeff_a * coeff_f * ds * delta_b;
        pb->data[nn] += coeff_b * (1 - coeff_f) * ds * delta_a;
        pb->data[nn + 1] += coeff_b * coeff_f * ds * delta_a;
        qa->data[lin] = a_avg;
        qb->data[lin] = b_avg;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------- movie monitor */

/* MovieMonitor<F>::update (movie_monitor.cu:17-45): every `interval` steps
 * the component over the monitor's box goes to frame t / interval. */
static void movie_record(xf_scene* s, int64_t t) {
  for (int32_t q = 0; q < s->n_movie; ++q) {
    xf_movie* m = &s->movie[q];
    if (t % m->interval != 0) continue;
    const int64_t frame = t / m->interval;
    if (frame >= m->data.shape[0]) continue;
    const xf_array* f = m->field < 3 ? efield(s, m->field) : hfield(s, m->field - 3);
    const xf_task* k = &m->task;
    const int64_t ny = k->y.end - k->y.start, nz = k->z.end - k->z.start;
    const int64_t nx = k->x.end - k->x.start;
    double* out = m->data.data + frame * nx * ny * nz;
    for (int64_t i = k->x.start; i < k->x.end; ++i)
      for (int64_t j = k->y.start; j < k->y.end; ++j)
        for (int64_t kk = k->z.start; kk < k->z.end; ++kk)
          out[((i - k->x.start) * ny + (j - k->y.start)) * nz + (kk - k->z.start)] = *at3(f, i, j, kk);
  }
}

/* ------------------------------------------------------- point sources */

/* Soft point sources: E_comp(i, j, k) += value[t], in registration order. */
static void point_sources(xf_scene* s, int64_t t) {
  for (int32_t q = 0; q < s->n_point; ++q) {
    const xf_point_source* p = &s->point[q];
    *at3(efield(s, p->comp), p->i, p->j, p->k) += p->value.data[t];
  }
}

/* ------------------------------------------------------- learned layer */

/* IEEE binary16 bits to float (exact). */
static float half_to_float(uint16_t h) {
  const int sign = h >> 15;
  const int ex = (h >> 10) & 0x1f;
  const int man = h & 0x3ff;
  float v;
  if (ex == 0)
    v = ldexpf((float)man, -24);
  else if (ex == 31)
    v = man ? NAN : INFINITY;
  else
    v = ldexpf((float)(man | 0x400), ex - 25);
  return sign ? -v : v;
}

/* The column MLP: mur + s * MLP(x / s) with s the power of two just above
 * max |x_i| (so the scaling is exact), 12 -> 32 ReLU -> 32 ReLU -> 1, in
 * the arithmetic the engine's matrix-core kernel defines:
 *   - the scaled inputs x / s and the layer-1 activations are rounded to
 *     fp16 (round to nearest even; the activations saturate at 65504);
 *   - each layer-1 unit is one 16-term matrix product (inputs 12..15 zero),
 *     each layer-2 unit two (hidden units 0..15, then 16..31), with the bias
 *     as the first accumulator; a product's 16 terms are added in two groups
 *     of 8 -- group 0 holds k-slots 0..7, group 1 k-slots 8..15 -- each group
 *     summed exactly with the accumulator and rounded once to fp32;
 *   - the layer-2 k-slot 8g + j of product p stands for hidden unit
 *     16p + 8(j / 4) + 4g + j % 4 (the row the layer-1 result lands in);
 *   - the output unit adds w3 . relu(h2) in two halves, g = 0 and 1, each over
 *     units 8(i / 4) + 4g + i % 4 for i = 0..15 in order (one rounding per
 *     product and per sum), then the halves, then b3.
 * The matrix cores' own adds of a group differ from an exact sum in the
 * last bits now and then (measured on gfx950: within 2 ulp), so the engine
 * agrees with this restatement to rounding, not bit for bit. */
static float f16_round(float v, int saturate) {
  if (v == 0.0f || !isfinite(v)) return v;
  double a = fabs((double)v);
  if (saturate && a > 65504.0) a = 65504.0;
  double q;
  if (a < 0x1p-14) {
    q = nearbyint(a * 0x1p24) * 0x1p-24;
  } else {
    int ex;
    (void)frexp(a, &ex); /* a in [2^(ex-1), 2^ex) */
    const double quantum = ldexp(1.0, ex - 1 - 10);
    q = nearbyint(a / quantum) * quantum;
  }
  if (q > 65504.0) q = INFINITY;
  return (float)(v < 0 ? -q : q);
}

/* every value of the MFMA chain is a multiple of 2^-48 (fp16 x fp16
 * products, fp16 biases and their fp32 roundings): exact in 128-bit fixed
 * point */
typedef __int128 lb_fix;

static lb_fix lb_to_fix(float v) { return (lb_fix)ldexp((double)v, 48); }

static float lb_round_fix(lb_fix n) {
  if (n == 0) return 0.0f;
  const int neg = n < 0;
  unsigned __int128 m = neg ? (unsigned __int128)(-n) : (unsigned __int128)n;
  int bl = 0;
  for (unsigned __int128 t = m; t; t >>= 1) ++bl;
  int sh = 0;
  if (bl > 24) {
    sh = bl - 24;
    const unsigned __int128 q = m >> sh, r = m - (q << sh), half = (unsigned __int128)1 << (sh - 1);
    m = q + ((r > half || (r == half && (q & 1))) ? 1 : 0);
  }
  const float f = ldexpf((float)(uint64_t)m, sh - 48);
  return neg ? -f : f;
}

/* c + sum_k a[k] b[k] over 16 k-slots, as two exact groups of 8 */
static float lb_mfma16(float c, const float* a, const float* b) {
  float acc = c;
  for (int g = 0; g < 2; ++g) {
    lb_fix s = lb_to_fix(acc);
    for (int j = 0; j < 8; ++j)
      s += (lb_fix)(int64_t)ldexp((double)a[8 * g + j], 24) * (lb_fix)(int64_t)ldexp((double)b[8 * g + j], 24);
    acc = lb_round_fix(s);
  }
  return acc;
}

static float lb_predict(const float* w, float mur_k, const float* x) {
  float m = 0.0f;
  for (int i = 0; i < XF_LB_IN; ++i) {
    const float ax = fabsf(x[i]);
    if (ax > m) m = ax;
  }
  const float mur = x[4] + mur_k * (x[0] - x[3]);
  if (m == 0.0f) return mur;
  int e;
  (void)frexpf(m, &e);
  float xs[16] = {0}, h1[XF_LB_HID], h2[XF_LB_HID];
  for (int i = 0; i < XF_LB_IN; ++i) xs[i] = f16_round(ldexpf(x[i], -e), 0);
  const float* W1 = w;
  const float* b1 = W1 + XF_LB_HID * XF_LB_IN;
  const float* W2 = b1 + XF_LB_HID;
  const float* b2 = W2 + XF_LB_HID * XF_LB_HID;
  const float* w3 = b2 + XF_LB_HID;
  const float b3 = w3[XF_LB_HID];
  for (int o = 0; o < XF_LB_HID; ++o) {
    float a[16] = {0};
    for (int i = 0; i < XF_LB_IN; ++i) a[i] = W1[o * XF_LB_IN + i];
    const float z = lb_mfma16(b1[o], a, xs);
    h1[o] = f16_round(z > 0.0f ? z : 0.0f, 1);
  }
  for (int o = 0; o < XF_LB_HID; ++o) {
    float acc = b2[o];
    for (int p = 0; p < 2; ++p) {
      float a[16], b[16];
      for (int g = 0; g < 2; ++g)
        for (int j = 0; j < 8; ++j) {
          const int u = 16 * p + 8 * (j / 4) + 4 * g + j % 4;
          a[8 * g + j] = W2[o * XF_LB_HID + u];
          b[8 * g + j] = h1[u];
        }
      acc = lb_mfma16(acc, a, b);
    }
    h2[o] = acc > 0.0f ? acc : 0.0f;
  }
  float half[2];
  for (int g = 0; g < 2; ++g) {
    float s = 0.0f;
    for (int i = 0; i < 16; ++i) {
      const int u = 8 * (i / 4) + 4 * g + i % 4;
      s = s + w3[u] * h2[u];
    }
    half[g] = s;
  }
  const float r = (half[0] + half[1]) + b3;
  return mur + ldexpf(r, e);
}

typedef struct lb_geom {
  int f, c, a, b, positive;
  int64_t n[3];
  int64_t na1, nb1;  /* n_a + 1, n_b + 1 */
  int64_t off;       /* history offset of the face */
} lb_geom;

static lb_geom lb_face(const xf_scene* s, int f) {
  lb_geom g;
  g.f = f;
  g.c = f / 2;
  g.positive = f & 1;
  g.a = (g.c + 1) % 3;
  g.b = (g.c + 2) % 3;
  g.n[0] = s->nx;
  g.n[1] = s->ny;
  g.n[2] = s->nz;
  g.na1 = g.n[g.a] + 1;
  g.nb1 = g.n[g.b] + 1;
  g.off = 0;
  for (int h = 0; h < f; ++h) {
    const int c = h / 2, a = (c + 1) % 3, b = (c + 2) % 3;
    g.off += 16 * (g.n[a] + 1) * (g.n[b] + 1);
  }
  return g;
}

/* in-plane index range of component cc (0: E_a, 1: E_b) */
static void lb_range(const lb_geom* g, int cc, int64_t inset, int64_t* p1, int64_t* q1) {
  *p1 = g->n[g->a] - inset - (cc == 0 ? 1 : 0);
  *q1 = g->n[g->b] - inset - (cc == 1 ? 1 : 0);
}

static double* lb_field(xf_scene* s, const lb_geom* g, int cc, int64_t inset, int layer,
                        int64_t p, int64_t q) {
  int64_t pos[3];
  pos[g->c] = g->positive ? g->n[g->c] - inset - layer : inset + layer;
  pos[g->a] = p;
  pos[g->b] = q;
  return at3(efield(s, cc == 0 ? g->a : g->b), pos[0], pos[1], pos[2]);
}

static double* lb_hist(xf_scene* s, const lb_geom* g, int par, int cc, int layer, int64_t p,
                       int64_t q) {
  return s->lb->history.data +
         g->off + ((((int64_t)par * 2 + cc) * 4 + layer) * g->na1 + p) * g->nb1 + q;
}

/* a node of component cc lies on another face of the layer: which (-1 if none) */
static int lb_other_face(const xf_scene* s, const lb_geom* g, int cc, int64_t p, int64_t q) {
  const int mask = s->lb->faces;
  if (cc == 0) { /* E_a: on the b faces at q = 0, n_b */
    if (q == 0 && (mask >> (2 * g->b)) & 1) return 2 * g->b;
    if (q == g->n[g->b] && (mask >> (2 * g->b + 1)) & 1) return 2 * g->b + 1;
  } else { /* E_b: on the a faces at p = 0, n_a */
    if (p == 0 && (mask >> (2 * g->a)) & 1) return 2 * g->a;
    if (p == g->n[g->a] && (mask >> (2 * g->a + 1)) & 1) return 2 * g->a + 1;
  }
  return -1;
}

static void lb_inputs(xf_scene* s, const lb_geom* g, int cc, int64_t inset, int64_t p,
                      int64_t q, int64_t p1, int64_t q1, int par, float* x) {
  for (int l = 1; l <= 3; ++l) x[l - 1] = (float)*lb_field(s, g, cc, inset, l, p, q);
  for (int l = 0; l < 4; ++l) x[3 + l] = (float)*lb_hist(s, g, par, cc, l, p, q);
  for (int l = 0; l < 3; ++l) x[7 + l] = (float)*lb_hist(s, g, par ^ 1, cc, l, p, q);
  /* in-plane 5-point Laplacians of u^t on layers 1 (x10) and 0 (x11); a
   * neighbour outside the column range counts as the centre value */
  for (int l = 1; l >= 0; --l) {
    const float u = x[3 + l];
    const float up = p < p1 ? (float)*lb_hist(s, g, par, cc, l, p + 1, q) : u;
    const float um = p > inset ? (float)*lb_hist(s, g, par, cc, l, p - 1, q) : u;
    const float vp = q < q1 ? (float)*lb_hist(s, g, par, cc, l, p, q + 1) : u;
    const float vm = q > inset ? (float)*lb_hist(s, g, par, cc, l, p, q - 1) : u;
    x[11 - l] = (((up + um) + (vp + vm)) - 4.0f * u);
  }
}

/* One step of the learned layer (after the E pass, its corrections and the
 * point sources of step t): APPLY writes the layer-0 nodes (faces' inner
 * nodes first, then the edge nodes each owned by its lower face), RECORD
 * stores samples; both then roll the history. */
static void lb_step(xf_scene* s, int64_t t) {
  xf_learned_boundary* lb = s->lb;
  const int64_t inset = lb->inset;
  const int par = (int)(t & 1);
  float w[XF_LB_NPARAM];
  if (lb->mode == XF_LB_APPLY)
    for (int i = 0; i < XF_LB_NPARAM; ++i) w[i] = half_to_float(lb->weights[i]);
  const float k = (float)lb->mur_k;
  const int record = lb->mode == XF_LB_RECORD && (t % lb->rec_every) == 0 &&
                     t / lb->rec_every < lb->rec_slots;
  int64_t col = 0, ncol = 0;
  for (int f = 0; f < 6; ++f) {
    if (!((lb->faces >> f) & 1)) continue;
    const lb_geom g = lb_face(s, f);
    for (int cc = 0; cc < 2; ++cc) {
      int64_t p1, q1;
      lb_range(&g, cc, inset, &p1, &q1);
      ncol += (p1 - inset + 1) * (q1 - inset + 1);
    }
  }
  for (int phase = 0; phase < 2; ++phase) {
    if (lb->mode == XF_LB_RECORD && phase == 1) break;
    col = 0;
    for (int f = 0; f < 6; ++f) {
      if (!((lb->faces >> f) & 1)) continue;
      const lb_geom g = lb_face(s, f);
      for (int cc = 0; cc < 2; ++cc) {
        int64_t p1, q1;
        lb_range(&g, cc, inset, &p1, &q1);
        for (int64_t p = inset; p <= p1; ++p)
          for (int64_t q = inset; q <= q1; ++q, ++col) {
            float x[XF_LB_IN];
            if (lb->mode == XF_LB_RECORD) {
              if (!record) continue;
              lb_inputs(s, &g, cc, inset, p, q, p1, q1, par, x);
              float* r = lb->rec + ((t / lb->rec_every) * ncol + col) * 13;
              for (int i = 0; i < XF_LB_IN; ++i) r[i] = x[i];
              r[12] = (float)*lb_field(s, &g, cc, inset, 0, p, q);
              continue;
            }
            const int other = lb_other_face(s, &g, cc, p, q);
            if (phase == 0 ? other >= 0 : !(other >= 0 && f < other)) continue;
            lb_inputs(s, &g, cc, inset, p, q, p1, q1, par, x);
            /* edge nodes (on two faces) take the first-order Mur update: the
             * network is trained on face interiors only */
            *lb_field(s, &g, cc, inset, 0, p, q) =
                phase == 0 ? (double)lb_predict(w, k, x) : (double)(x[4] + k * (x[0] - x[3]));
          }
      }
    }
  }
  if (record && t / lb->rec_every + 1 > lb->rec_count) lb->rec_count = t / lb->rec_every + 1;
  /* history: u^{t+1} of layers 0..3 into the parity buffer of t + 1 */
  for (int f = 0; f < 6; ++f) {
    if (!((lb->faces >> f) & 1)) continue;
    const lb_geom g = lb_face(s, f);
    for (int cc = 0; cc < 2; ++cc) {
      int64_t p1, q1;
      lb_range(&g, cc, inset, &p1, &q1);
      for (int64_t p = inset; p <= p1; ++p)
        for (int64_t q = inset; q <= q1; ++q)
          for (int l = 0; l < 4; ++l)
            *lb_hist(s, &g, par ^ 1, cc, l, p, q) =
                (double)(float)*lb_field(s, &g, cc, inset, l, p, q);
    }
  }
}

/* --------------------------------------------------------------- the loop */

static char g_err[512];

const char* xfo_last_error(void) { return g_err; }

int xfo_run(xf_scene* s, int64_t n_steps) {
  g_err[0] = 0;
  if (!s || s->abi_version != XF_ABI_VERSION) {
    snprintf(g_err, sizeof g_err, "bad scene");
    return 1;
  }
  for (int64_t n = 0; n < n_steps; ++n) {
    const int64_t t = s->current_step;
    /* updateH: BasicUpdator::updateH (basic_updator.cuh:16-32) */
    yee_update(s, 0, 0);
    yee_update(s, 0, 1);
    yee_update(s, 0, 2);
    /* correctH (domain_hd.cu:51-55) */
    if (s->tfsf) tfsf_correct(s, 0, t);
    for (int32_t p = 0; p < s->n_pml; ++p) pml_correct(s, &s->pml[p], 0);
    /* updateE */
    if (s->ade && s->ade->kind != XF_ADE_NONE) {
      ade_update_e(s, 0);
      ade_update_e(s, 1);
      ade_update_e(s, 2);
    } else if (s->dimension == 2) {
      yee_update(s, 1, 2); /* BasicUpdatorTE::updateE */
    } else {
      yee_update(s, 1, 0);
      yee_update(s, 1, 1);
      yee_update(s, 1, 2);
    }
    /* correctE */
    if (s->tfsf) tfsf_correct(s, 1, t);
    for (int32_t p = 0; p < s->n_pml; ++p) pml_correct(s, &s->pml[p], 1);
    /* additions without a reference counterpart */
    point_sources(s, t);
    if (s->lb && s->lb->mode != XF_LB_OFF) lb_step(s, t);
    /* record (domain_hd.cu:65-77): monitors, then NF2FF */
    movie_record(s, t);
    for (int32_t q = 0; q < s->n_nf2ff; ++q) {
      const xf_nf2ff* nf = &s->nf2ff[q];
      for (int32_t f = 0; f < nf->n_freq; ++f)
        for (int face = 0; face < 6; ++face)
          nf2ff_face(s, nf, &nf->freq[f], face, t);
    }
    /* NF2FFTimeDoaminAgency::update: faces XN..ZP, E then H (:192-223) */
    for (int32_t q = 0; q < s->n_nf2ff_td; ++q) {
      if (s->dimension != 3) {
        snprintf(g_err, sizeof g_err, "nf2ff_td needs a 3-D grid");
        return 1;
      }
      for (int face = 0; face < 6; ++face)
        for (int e = 1; e >= 0; --e)
          if (nf2ff_td_face(s, &s->nf2ff_td[q], face, e, t)) {
            snprintf(g_err, sizeof g_err, "nf2ff_td: time bin outside the potential arrays at step %lld",
                     (long long)t);
            return 1;
          }
    }
    /* nextStep */
    s->current_step += 1;
  }
  return 0;
}

/* The learned layer's column MLP on its own (test hook): n columns of 12
 * inputs -> n predictions, weights as in xf_learned_boundary. */
int xfo_lb_predict(const uint16_t* weights, double mur_k, const float* x, int64_t n,
                   float* out) {
  float w[XF_LB_NPARAM];
  for (int i = 0; i < XF_LB_NPARAM; ++i) w[i] = half_to_float(weights[i]);
  for (int64_t c = 0; c < n; ++c) out[c] = lb_predict(w, (float)mur_k, x + c * XF_LB_IN);
  return 0;
}

/* The matrix-core product model on its own (test hook): out[r] =
 * c[r] + sum_k a[r][k] b[r][k] over 16 k-slots (fp16 values given as
 * floats), two exact groups of 8 with one fp32 rounding each. */
int xfo_lb_mfma16(const float* c, const float* a, const float* b, int64_t n, float* out) {
  for (int64_t r = 0; r < n; ++r) out[r] = lb_mfma16(c[r], a + 16 * r, b + 16 * r);
  return 0;
}

int xfo_num_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void xfo_set_num_threads(int n) {
#ifdef _OPENMP
  extern void omp_set_num_threads(int);
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* One Yee component update in the flat linear-index form of the engine's
 * reference path (update_scheme.cuh:141-197). */
int xfo_yee_component(xf_scene* s, int is_e, int xyz) {
  yee_update(s, is_e, xyz);
  return 0;
}

/* Brute-force form of one Yee component update, written as the reference's
 * triple-loop update (update_scheme.cuh:65-130) instead of the flat
 * linear-index loop; tests check that both forms agree. */
int xfo_yee_component_triple_loop(xf_scene* s, int is_e, int xyz) {
  const int a_axis = (xyz + 1) % 3, b_axis = (xyz + 2) % 3;
  const xf_array* f = is_e ? efield(s, xyz) : hfield(s, xyz);
  const xf_array* fa = is_e ? hfield(s, a_axis) : efield(s, a_axis);
  const xf_array* fb = is_e ? hfield(s, b_axis) : efield(s, b_axis);
  const int base = (is_e ? 0 : 9) + 3 * xyz;
  int64_t is = 0, js = 0, ks = 0;
  /* The triple-loop form has no a==0||b==0 guard; start E loops at 1 on the
   * two tangential axes instead. */
  if (is_e) {
    if (a_axis == 0 || b_axis == 0) is = 1;
    if (a_axis == 1 || b_axis == 1) js = 1;
    if (a_axis == 2 || b_axis == 2) ks = 1;
  }
  for (int64_t i = is; i < s->nx; ++i)
    for (int64_t j = js; j < s->ny; ++j)
      for (int64_t k = ks; k < s->nz; ++k) {
        int64_t di[3] = {0, 0, 0};
        const int64_t off = is_e ? -1 : 1;
        /* neighbour of field_a along axis b, of field_b along axis a */
        di[b_axis] = off;
        const double fa_q = *at3(fa, i + di[0], j + di[1], k + di[2]);
        di[b_axis] = 0;
        di[a_axis] = off;
        const double fb_q = *at3(fb, i + di[0], j + di[1], k + di[2]);
        const double cf = coef(s, base, i, j, k);
        const double ca = coef(s, base + 1, i, j, k);
        const double cb = coef(s, base + 2, i, j, k);
        const double fv = *at3(f, i, j, k);
        if (is_e)
          *at3(f, i, j, k) = e_next(cf, fv, ca, *at3(fa, i, j, k), fa_q, cb,
                                    *at3(fb, i, j, k), fb_q);
        else
          *at3(f, i, j, k) = h_next(cf, fv, ca, *at3(fa, i, j, k), fa_q, cb,
                                    *at3(fb, i, j, k), fb_q);
      }
  return 0;
}
