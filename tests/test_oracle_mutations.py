"""The composite-operator pins must have teeth: each plausible mistake below
(a dropped or flipped term, a wrong stagger's metric factor, an index off
by one, a transposed operand), planted in a copy of the C oracle built on
the CPU, must make the named test of tests/test_oracle_compositions.py
(or of the file a fourth entry names) fail.  The repository's oracle
itself is never modified."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "pcom_oracle.c")

PINS = "test_oracle_pins.py"

# name: (statement in oracle/pcom_oracle.c, mistaken statement, test that
# must fail[, test file]); the first occurrence of the statement is changed
MUTANTS = {
    # smag
    "smag_shear_metric_sign": ("u[I3(i, j, k)] * tn / ARAD;", "u[I3(i, j, k)] * tn / ARAD * -1.0;",
                               "smag_closed_form", PINS),
    "smag_factor_of_four": ("double ds = 4 * (v[I3(i + 1, j, k)] - v[I3(i - 1, j, k)]) / dx +",
                            "double ds = 2 * (v[I3(i + 1, j, k)] - v[I3(i - 1, j, k)]) / dx +",
                            "smag_closed_form", PINS),
    "smag_tn_of_T_row": ("double tn = o->tn[3][I2(i, j)];", "double tn = o->tn[1][I2(i, j)];",
                         "smag_closed_form", PINS),
    "smag_cap_divisor": ("(double)(4 * 8 * o->dtbc)", "(double)(4 * 4 * o->dtbc)", "smag_cap", PINS),
    "smag_last_row": ("for (int j = 2; j <= NJ - 1; ++j)", "for (int j = 2; j <= NJ - 2; ++j)",
                      "smag_cap", PINS),
    # den_prho
    "prho_interface_pressure": ("double p = o->chc[I2(i, j)] * o->pr2[k - 1];",
                                "double p = o->chc[I2(i, j)] * o->pr1[k - 1];", "den_prho", PINS),
    "prho_no_chc": ("double p = o->chc[I2(i, j)] * o->pr2[k - 1];", "double p = o->pr2[k - 1];",
                    "den_prho", PINS),
    "prho_one_sided_divisor": ("o->prho1[I3(i, j, k)] = (ra - rb) / (2 * DELTAT);",
                               "o->prho1[I3(i, j, k)] = (ra - rb) / DELTAT;", "den_prho", PINS),
    "prho_components_swapped": ("o->prho2[I3(i, j, k)] = (ra - rb) / (2 * DELTAS);",
                                "o->prho1[I3(i, j, k)] = (ra - rb) / (2 * DELTAS);", "den_prho", PINS),
    # den_alpha
    "alpha_interface_pressure": ("ch[I2(i, j)] * o->pr2[k - 1]);", "ch[I2(i, j)] * o->pr1[k - 1]);",
                                 "den_alpha", PINS),
    "alpha_no_ch": ("ch[I2(i, j)] * o->pr2[k - 1]);", "o->pr2[k - 1]);", "den_alpha", PINS),
    # int_bnd
    "bnd_century_leap": ("if (y % 100 == 0) { if (y % 400 == 0) d = 366; }", "if (y % 100 == 0) { d = 366; }",
                         "int_bnd_interpolates", PINS),
    "bnd_january_wrap": ("m2 = m; m1 = m - 1; if (m == 1) m1 = 12;", "m2 = m; m1 = m - 1; if (m == 1) m1 = 1;",
                         "int_bnd_interpolates", PINS),
    "bnd_length_of_later_month": ("days = days_month(y, m1);", "days = days_month(y, m2);",
                                  "int_bnd_interpolates", PINS),
    "bnd_weight_in_double": ("double fd = (double)f;",
                             "double fd = (d < 15) ? ((double)days - 15.0 + (double)d) / (double)days "
                             ": ((double)d - 15.0) / (double)days; (void)f;",
                             "int_bnd_interpolates", PINS),
    # int_readyc
    "readyc_pa_force_sign": ("double x = -o->auc[q] - sch[p] * A0 * o->pax[p] + o->fx[q];",
                             "double x = -o->auc[q] + sch[p] * A0 * o->pax[p] + o->fx[q];",
                             "pressure_force_at_rest", PINS),
    "readyc_pa_force_no_sch": ("double x = -o->auc[q] - sch[p] * A0 * o->pax[p] + o->fx[q];",
                               "double x = -o->auc[q] - A0 * o->pax[p] + o->fx[q];",
                               "pressure_force_at_rest", PINS),
    "readyc_pa_gradient_staggers": ("gra(o, wkx, wky, o->bnd[4], 1, 1, 2, 4);",
                                    "gra(o, wkx, wky, o->bnd[4], 1, 1, 4, 2);",
                                    "pressure_force_at_rest", PINS),
    "readyc_bootstrap_on_restart": ("if (o->rst != 1) {", "if (o->rst != 2) {", "adams_bashforth_history", PINS),
    "readyc_bootstrap_second_step": ("copy3(o, o->aupp, o->aup); copy3(o, o->avpp, o->avp);",
                                     "copy3(o, o->aupp, o->auc); copy3(o, o->avpp, o->avc);",
                                     "adams_bashforth_history", PINS),
    "readyc_shift_order": ("copy3(o, o->aupp, o->aup);\n  copy3(o, o->avpp, o->avp);\n  copy3(o, o->aup, o->auc);\n"
                           "  copy3(o, o->avp, o->avc);",
                           "copy3(o, o->aup, o->auc);\n  copy3(o, o->avp, o->avc);\n  copy3(o, o->aupp, o->aup);\n"
                           "  copy3(o, o->avpp, o->avp);",
                           "adams_bashforth_history", PINS),
    # int_ts accumulators
    "ts_acr_pressure_of_mch": ("den_alpha(o, wk, o->tp, o->sp, o->chc, o->msk_t);",
                               "den_alpha(o, wk, o->tp, o->sp, o->mch, o->msk_t);", "density_accumulator", PINS),
    "ts_acr_specific_volume": ("o->acr[q] = o->acr[q] + 1.0 / wk[q];", "o->acr[q] = o->acr[q] + wk[q];",
                               "density_accumulator", PINS),
    "ts_acw_divisor": ("o->wm[q] * o->alpha[q] / (GRAV * o->mch2[p]);", "o->wm[q] * o->alpha[q] / (GRAV * o->mch[p]);",
                       "vertical_velocity_accumulator", PINS),
    "ts_acw_interface": ("o->acw[q] = o->acw[q] + o->wm[q] * o->alpha[q]",
                         "o->acw[q] = o->acw[q] + o->wm[q + n2] * o->alpha[q]",
                         "vertical_velocity_accumulator", PINS),
    # op_adv
    "adv_half_correction": ("ans[q] = ans[q] - 0.5 * var[q] * ((m1[q] + m2[q])",
                            "ans[q] = ans[q] - 0.25 * var[q] * ((m1[q] + m2[q])", "skew_symmetric"),
    "adv_V_metric": ("      wka[q] = wka[q] * o->rh[3][p];", "      wka[q] = wka[q] * o->rh[1][p];",
                     "nondivergent_flow"),
    "adv_vertical_sign": ("  for (size_t q = 0; q < n3; ++q) ans[q] = ans[q] + wka[q];\n  p1(o, m1, u",
                          "  for (size_t q = 0; q < n3; ++q) ans[q] = ans[q] - wka[q];\n  p1(o, m1, u",
                          "skew_symmetric"),
    # op_adv_ts
    "advts_V_metric": ("um2[q] = o->vc[q] * sch[p] * o->rh[3][p];", "um2[q] = o->vc[q] * sch[p] * o->rh[1][p];",
                       "adv_ts_of_coordinates"),
    "advts_transposed": ("ter(o, wka, um1, NK, 3, 2, 0.0);", "ter(o, wka, um2, NK, 3, 2, 0.0);",
                         "adv_ts_of_coordinates"),
    "advts_dpr_index": ("wka[p + n2 * k] = wka[p + n2 * k] / o->dpr2[k];",
                        "wka[p + n2 * k] = wka[p + n2 * k] / o->dpr2[k > 0 ? k - 1 : 0];",
                        "summation_by_parts"),
    # op_dif
    "dif_x_metric": ("mx[p + n2 * k] = mx[p + n2 * k] / (ARAD * o->rh[2][p]);",
                     "mx[p + n2 * k] = mx[p + n2 * k] / (ARAD * o->rh[4][p]);", "dif_closed_form"),
    "dif_ch_faces": ("  ter(o, cha, ch, 1, 1, 2, 0.0);\n  ter(o, chb, ch, 1, 1, 4, 0.0);",
                     "  ter(o, cha, ch, 1, 1, 4, 0.0);\n  ter(o, chb, ch, 1, 1, 2, 0.0);", "dif_closed_form"),
    "dif_vertical_sign": ("wkp[q] = GRAV * RHO0 * o->kh[q] * wkp[q];", "wkp[q] = -GRAV * RHO0 * o->kh[q] * wkp[q];",
                          "vertical_part_dissipates"),
    # fri_r3d / op_fri
    "fri_cv1_sign": ("ans[q] = ans[q] + o->am[q] * o->cv1[p] * vp[q];",
                     "ans[q] = ans[q] - o->am[q] * o->cv1[p] * vp[q];", "rigid_rotations"),
    "fri_cv2_sign": ("ans[q] = ans[q] + o->am[q] * o->cv2[p] * sch[p] * wka[q];",
                     "ans[q] = ans[q] - o->am[q] * o->cv2[p] * sch[p] * wka[q];", "rigid_rotations"),
    "fri_cross_sign": ("fri_r3d(o, o->fx, sch, u, o->uc, mv, o->bnd[0]);",
                       "fri_r3d(o, o->fx, sch, u, o->uc, v, o->bnd[0]);", "rigid_rotations"),
    "fri_drag_level": ("wkp[I3(i, j, k + 1)] = vv * GRAV", "wkp[I3(i, j, k)] = vv * GRAV", "wind_minus_drag"),
    "fri_km_weight": ("wkp[q] = GRAV * RHO0 * o->km[q] * wkp[q] / (sch[p] * sch[p]);",
                      "wkp[q] = GRAV * RHO0 * o->km[q] * wkp[q] / sch[p];", "linear_shear"),
    # upwelling
    "upw_dpr_index": ("o->wm[q] = o->wm[qm] + o->dpr2[k - 2] * (wk2[p] + wk3[qm]);",
                      "o->wm[q] = o->wm[qm] + o->dpr2[k - 1] * (wk2[p] + wk3[qm]);", "one_level_flow"),
    "upw_div_level": ("o->wm[q] = o->wm[qm] + o->dpr2[k - 2] * (wk2[p] + wk3[qm]);",
                      "o->wm[q] = o->wm[qm] + o->dpr2[k - 2] * (wk2[p] + wk3[q]);", "one_level_flow"),
    "upw_mean_sign": ("for (size_t p = 0; p < n2; ++p) wk2[p] = -wk2[p];", "", "barotropic_flow"),
    # int_pgra
    "pgra_cshift_direction": ("      else js = (j == 1) ? NJ : j - 1;", "      else js = (j == NJ) ? 1 : j + 1;",
                              "constant_specific_volume"),
    "pgra_interface_pressure": ("      wk[q] = wkb[q] + o->pr2[k] * wkc[q];\n    }\n  vint_face(o, wk, o->bphi[0]",
                                "      wk[q] = wkb[q] + o->pr1[k] * wkc[q];\n    }\n  vint_face(o, wk, o->bphi[0]",
                                "constant_specific_volume"),
    "pgra_gradients_swapped": ("  vint_face(o, wkx, o->bgra[0], o->bgra[1], 2);\n  vint_face(o, wky, o->bgra[2], o->bgra[3], 1);",
                               "  vint_face(o, wky, o->bgra[0], o->bgra[1], 2);\n  vint_face(o, wkx, o->bgra[2], o->bgra[3], 1);",
                               "linear_specific_volume"),
    # int_trop
    "trop_topography_sign": ("pg1[p] = (gi1[p] + ci1[p] + o->graphihx[p]) * su * sch[p];",
                             "pg1[p] = (gi1[p] + ci1[p] - o->graphihx[p]) * su * sch[p];", "at_rest"),
    "trop_face_pairing": ("ci1[I2(i, j)] = 0.5 * (wk[I2(i, j)] * o->bgra[0][I2(i, j)] + wk[I2(i, j + 1)] * o->bgra[1][I2(i, j + 1)]);",
                          "ci1[I2(i, j)] = 0.5 * (wk[I2(i, j)] * o->bgra[0][I2(i, j)] + wk[I2(i, j + 1)] * o->bgra[0][I2(i, j + 1)]);",
                          "at_rest"),
    "trop_free_surface_sign": ("pg2[p] = (gi2[p] + ci2[p] + o->graphihy[p]) * su * sch[p];",
                               "pg2[p] = (-gi2[p] + ci2[p] + o->graphihy[p]) * su * sch[p];", "geostrophic"),
    "trop_coriolis_x": ("double x = 0.5 * o->cor[p] * bt;", "double x = o->cor[p] * bt;", "inertial"),
    "trop_no_half_step": ("o->buc[p] = 0.5 * (o->buc[p] + o->bup[p]);", "", "inertial"),
    "trop_mean_count": ("o->buc[p] = upb1[p] / (double)(nstep + 1);", "o->buc[p] = upb1[p] / (double)(nstep);",
                        "inertial"),
    # int_clin
    "clin_no_barotropic": ("o->uc[q] = o->uc[q] - wm1[p] + o->buc[p];", "o->uc[q] = o->uc[q] - wm1[p];",
                           "depth_mean_is_barotropic"),
    "clin_full_step_coriolis": ("double c = 0.5 * o->cor[p] * bc;", "double c = o->cor[p] * bc;",
                                "crank_nicolson"),
    "clin_rotation_sense": ("pu1[p] = x1 + c * x2;\n      pu2[p] = x2 - c * x1;",
                            "pu1[p] = x1 - c * x2;\n      pu2[p] = x2 + c * x1;", "crank_nicolson"),
    "clin_ab3_history": ("c_tp * o->aup[q] + c_tpp * o->aupp[q]);", "c_tp * o->aupp[q] + c_tpp * o->aup[q]);",
                         "ab3"),
    # int_ssh
    "ssh_column_sign": ("o->ssh[p] = o->phih_t[p] / GRAV + o->ssh[p];", "o->ssh[p] = o->phih_t[p] / GRAV - o->ssh[p];",
                        "reference_stratification"),
    "ssh_interface_dpr": ("s += o->alpha[I3(i, j, k)] * o->dpr2[k - 1] * o->mch[I2(i, j)]",
                          "s += o->alpha[I3(i, j, k)] * o->dpr1[k - 1] * o->mch[I2(i, j)]",
                          "reference_stratification"),
    # int_ts
    "ts_restoring_layer": ("double pb = GAMMA_T * (o->bnd[2][p] - o->tc[p]) * (double)o->msk_t[p] / o->dpr2[0];",
                           "double pb = GAMMA_T * (o->bnd[2][p] - o->tc[p]) * (double)o->msk_t[p] / o->dpr2[1];",
                           "restoring_budget"),
    "ts_no_mch": ("o->tc[q] = o->tp[q] + pt1[q] * bc * (double)o->msk_t[q] / o->mch[p];",
                  "o->tc[q] = o->tp[q] + pt1[q] * bc * (double)o->msk_t[q];", "restoring_budget"),
    # convect
    "convect_not_called": ("  convect(o);\n  copy3(o, o->tp, o->tc);", "  copy3(o, o->tp, o->tc);",
                           "convect_leaves_unstable", PINS),
    "convect_rhob_own_level": ("rhob[k] = p1c[k + 1] * t[k] + p2c[k + 1] * s[k];",
                               "rhob[k] = p1c[k] * t[k] + p2c[k] * s[k];", "convect_leaves_unstable", PINS),
}

_RUN = """
import sys
sys.path.insert(0, {root!r})
import oracle.oracle as oo
oo._HERE = {here!r}
import pytest
sys.exit(pytest.main(["-q", "-p", "no:cacheprovider", "--tb=no", "-k", {k!r},
                      {tests!r}]))
"""


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_is_caught(name, tmp_path):
    good, bad, test = MUTANTS[name][:3]
    tfile = MUTANTS[name][3] if len(MUTANTS[name]) > 3 else "test_oracle_compositions.py"
    src = open(SRC).read()
    assert src.count(good) >= 1, f"{name}: statement no longer in the oracle"
    (tmp_path / "pcom_oracle.c").write_text(src.replace(good, bad, 1))
    lib = tmp_path / "liboracle.so"
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-ffp-contract=off", "-std=c11",
                           "-I", os.path.join(ROOT, "oracle"), "-shared", "-o", str(lib),
                           str(tmp_path / "pcom_oracle.c"), "-lm"])
    tests = os.path.join(ROOT, "tests", tfile)
    code = _RUN.format(root=ROOT, here=str(tmp_path), k=test, tests=tests)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert "passed" in p.stdout or "failed" in p.stdout, p.stdout + p.stderr
    assert p.returncode == 1 and "failed" in p.stdout, f"{name} not caught by {test}:\n{p.stdout[-500:]}"
