"""The oracle pins catch plausible mistakes.

Each case copies oracle/mimadv_oracle.c, plants one mistake of the kind a
restatement of the reference could make (a dropped factor, a shifted index,
a wrong level metric, a swapped branch, a missing guard), builds the copy
with the oracle's own compiler flags into a temporary directory, and runs
the known-answer check that covers that code against it: the check must
fail.  The unmodified oracle passes every check (test_oracle_known_answers).
"""
import os
import shutil
import subprocess

import pytest

import test_oracle_known_answers as ka
import test_oracle_positive as pos

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "mimadv_oracle.c")
HDR = os.path.join(ROOT, "oracle", "mimadv_oracle.h")
CFLAGS = ["-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared"]

# (name, text in the oracle, planted replacement, check, check arguments)
MUTANTS = [
    # positivity-preserving stage
    ("positive donor of a positive x flux", "? th[IDX(g, WX(f), j, k)]", "? th[IDX(g, WX((f) + 1), j, k)]",
     pos.check_positive_single_cell, ()),
    ("positive negative cells donate", "double qp = q[id] > 0.0 ? q[id] : 0.0;", "double qp = q[id];",
     pos.check_positive_invariants, ()),
    ("positive y outflow dropped", "if (three_d) o = o + (POS(fy[id])", "if (0) o = o + (POS(fy[id])",
     pos.check_positive_invariants, ()),
    ("positive z outflow metric", "o = o + (POS(fzt) + NEG(fzb)) * L(g->dz1w, k);",
     "o = o + (POS(fzt) + NEG(fzb)) * L(g->dz1w, k < nz ? k + 1 : k);", pos.check_positive_invariants, ()),
    ("positive divergence order",
     "A = A - X;\n                if (three_d) A = A - Y;\n                A = A - Z;",
     "A = A - Z;\n                A = A - X;\n                if (three_d) A = A - Y;", pos.check_positive_identity, ()),
    ("positive theta without the outflow test", "th[id] = o > qp ? (qp / o) * (1.0 - 0x1p-40) : 1.0;",
     "th[id] = (qp / o) * (1.0 - 0x1p-40);", pos.check_positive_identity, ()),
    ("positive theta without the margin", "th[id] = o > qp ? (qp / o) * (1.0 - 0x1p-40) : 1.0;",
     "th[id] = o > qp ? (qp / o) : 1.0;", pos.check_positive_single_cell, ()),
    # momentum artificial viscosity
    ("avisc self mu index", "double mu = fabs(L(ut, i)) * d1;", "double mu = fabs(L(ut, i + 1)) * d1;",
     ka.check_lavisc_site, ("ux",)),
    ("avisc self mu in y", "double mu = fabs(L(ut, i)) * d1;", "double mu = fabs(L(ut, i)) * d1 * 0.5;",
     ka.check_lavisc_site, ("vy",)),
    ("avisc cross corner average",
     "double mu = 0.25 * fabs(L(t1, i) + L(t2, i) + L(t1, i + 1) + L(t2, i + 1)) * d1;",
     "double mu = 0.25 * fabs(L(t1, i) + L(t2, i) + L(t1, i - 1) + L(t2, i - 1)) * d1;",
     ka.check_lavisc_site, ("uy",)),
    ("avisc cross abs of terms",
     "double mu = 0.25 * fabs(L(t1, i) + L(t2, i) + L(t1, i + 1) + L(t2, i + 1)) * d1;",
     "double mu = 0.25 * (fabs(L(t1, i)) + fabs(L(t2, i)) + fabs(L(t1, i + 1)) + fabs(L(t2, i + 1))) * d1;",
     ka.check_lavisc_site, ("vx",)),
    ("avisc z level metric",
     "double mu = 0.25 * fabs(L(t1, k) + L(t2, k) + L(t1, k + 1) + L(t2, k + 1)) * L(dz1w, k);",
     "double mu = 0.25 * fabs(L(t1, k) + L(t2, k) + L(t1, k + 1) + L(t2, k + 1)) * L(dz1w, k + 1);",
     ka.check_lavisc_site, ("uz",)),
    ("avisc z range", "for (int k = 3; k <= nz - 2; ++k) {\n            double mu = 0.25",
     "for (int k = 3; k <= nz - 1; ++k) {\n            double mu = 0.25", ka.check_lavisc_site, ("vz",)),
    ("avisc w horizontal average",
     "double mu = 0.25 * fabs(T2(i, k) + T2(i, k - 1) + T2(i + 1, k) + T2(i + 1, k - 1)) * d1;",
     "double mu = 0.25 * fabs(T2(i, k) + T2(i, k + 1) + T2(i + 1, k) + T2(i + 1, k + 1)) * d1;",
     ka.check_lavisc_site, ("wx",)),
    ("avisc w horizontal level range", "for (int k = 2; k <= nz; ++k)\n            for (int i = i0; i <= i1; ++i) {\n"
     "                double mu", "for (int k = 3; k <= nz; ++k)\n            for (int i = i0; i <= i1; ++i) {\n"
     "                double mu", ka.check_lavisc_site, ("wy",)),
    ("avisc wz metric", "double mu = fabs(L(wt, k)) * L(dz1, k - 1);", "double mu = fabs(L(wt, k)) * L(dz1, k);",
     ka.check_lavisc_site, ("wz",)),
    ("avisc fourth difference", "- 6.0 * L(u, i) + 4.0 * L(u, i - 1)", "- 5.0 * L(u, i) + 4.0 * L(u, i - 1)",
     ka.check_lavisc_site, ("ux",)),
    # lim_tol gates
    ("muscl gate without dx", "if (d * fabs(slope_m) > lt * fabs(xx) + eps || d * fabs(slope_p2) > lt * fabs(xx) + eps)",
     "if (fabs(slope_m) > lt * fabs(xx) + eps || fabs(slope_p2) > lt * fabs(xx) + eps)", ka.check_muscl_gates, ()),
    ("muscl gate on the right cell",
     "if (d * fabs(slope_m) > lt * fabs(xx) + eps || d * fabs(slope_p2) > lt * fabs(xx) + eps)",
     "if (d * fabs(slope_m) > lt * fabs(xx) + eps || d * fabs(slope_p2) > lt * fabs(L(x, i + 1)) + eps)",
     ka.check_muscl_gates, ()),
    ("muscl r guard in index units",
     "if (fabs(ffim1) > lt * fabs(xx) && fabs(ffip1) < fabs(ffim1) / lt) rp1 = ffip1 / ffim1;",
     "if (50.0 * fabs(ffim1) > lt * fabs(xx) && fabs(ffip1) < fabs(ffim1) / lt) rp1 = ffip1 / ffim1;",
     ka.check_muscl_gates, ()),
    ("muscl second r guard dropped",
     "if (fabs(ffim1) > lt * fabs(xx) && fabs(ffip1) < fabs(ffim1) / lt) rp1 = ffip1 / ffim1;",
     "if (fabs(ffim1) > lt * fabs(xx)) rp1 = ffip1 / ffim1;", ka.check_muscl_gates, ()),
    ("quick rp11 guard on ffim1", "if (fabs(ffip1) > lt * fabs(phic) + eps) {\n        rp11",
     "if (fabs(ffim1) > lt * fabs(phic) + eps) {\n        rp11", ka.check_quick_gates, ()),
    ("quick clip bound", "*phil = dmax(0.0, dmin(2.0 * rp1, dmin(*phil, 2.0)));",
     "*phil = dmax(0.0, dmin(1.0 * rp1, dmin(*phil, 2.0)));", ka.check_quick_gates, ()),
    ("quick clip dropped", "if (fabs(ffim1) > lt * fabs(phic) + eps || fabs(ffip2) > lt * fabs(phic) + eps) {",
     "if (0) {", ka.check_quick_gates, ()),
    ("lw gate and", "if (fabs(ffim1) > lt * fabs(xx) + eps || fabs(ffip2) > lt * fabs(xx) + eps) {\n            double phir, phil;\n            lw_limiter",
     "if (fabs(ffim1) > lt * fabs(xx) + eps && fabs(ffip2) > lt * fabs(xx) + eps) {\n            double phir, phil;\n            lw_limiter",
     ka.check_lw_gates, ()),
    # PPM tile edges and face fixes
    ("ppm left closure face", "xmin = 0.5 * (L(x, ii1 + 1) + L(x, ii1));", "xmin = 0.5 * (L(x, ii1) + L(x, ii1 - 1));",
     ka.check_ppm_tile_edges, ()),
    ("ppm right closure face", "xplus = L(x, ii2 + 1) + 0.5 * ffip1;", "xplus = L(x, ii2 + 1) + 0.5 * ffim1;",
     ka.check_ppm_tile_edges, ()),
    ("ppm right closure cell skipped", "L(fl, ii2 + 1) = ppm_fl(&l, muc);\n    /* non-periodic",
     "L(fl, ii2 + 1) = ppm_fl(&l, mup);\n    /* non-periodic", ka.check_ppm_tile_edges, ()),
    ("ppm face fix to the far value", "if (fabs(r.phir - xx) < fabs(r.phir - xp)) r.phir = xx;\n            else r.phir = xp;",
     "if (fabs(r.phir - xx) < fabs(r.phir - xp)) r.phir = xp;\n            else r.phir = xx;",
     ka.check_ppm_face_fix, ()),
    ("ppm left positivity dropped", "if (r.phil < 0.0 || ((xx - r.phil) * (r.phil - xm) < 0.0 &&",
     "if (((xx - r.phil) * (r.phil - xm) < 0.0 &&", ka.check_ppm_face_fix, ()),
    ("ppm left fix to xx", "if (fabs(r.phil - xx) < fabs(r.phil - xm)) r.phil = xx;\n            else r.phil = xm;",
     "r.phil = xx;", ka.check_ppm_face_fix, ()),
    # integ_mom and its sponge
    ("sponge pi", "const double pi = 3.14159265, taud = 0.25;", "const double pi = 3.141592653589793, taud = 0.25;",
     ka.check_integ_mom, ()),
    ("sponge a_lm", "tauw = 1.0 / (1.0 + a_lm * dt * tauw);", "tauw = 1.0 / (1.0 + dt * tauw);",
     ka.check_integ_mom, ()),
    ("sponge first level wins", "            if (L(z0, k) >= zdamp) {\n                double c",
     "            if (L(z0, k) >= zdamp && tauw == 1.0) {\n                double c", ka.check_integ_mom, ()),
    ("integ_mom b_lm", "    double tw = orc_sponge_factor(g->nz, sorder, im, dt, z0, zmax, zdamp, sponge);",
     "    b_lm = 0.5 * b_lm;\n    double tw = orc_sponge_factor(g->nz, sorder, im, dt, z0, zmax, zdamp, sponge);",
     ka.check_integ_mom, ()),
    # WENO5 smoothness indicators
    ("weno5 beta0 second term",
     "0.25 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c)", "0.25 * (a - 4.0 * b + c) * (a - 4.0 * b + c)",
     ka.check_weno5_indicators, ()),
    ("weno5 beta1 coefficient", "double b1 = 13.0 / 12.0 *", "double b1 = 1.0 / 12.0 *",
     ka.check_weno5_indicators, ()),
    ("weno5 beta2 stencil", "0.25 * (3.0 * c - 4.0 * d + e) * (3.0 * c - 4.0 * d + e)",
     "0.25 * (3.0 * c - 4.0 * d + e) * (c - 4.0 * d + 3.0 * e)", ka.check_weno5_indicators, ()),
    ("weno5 z downward stencil", "weno5_left(XZ(k + 3), XZ(k + 2), XZ(k + 1), XZ(k), XZ(k - 1));",
     "weno5_left(XZ(k - 1), XZ(k), XZ(k + 1), XZ(k + 2), XZ(k + 3));", ka.check_weno5_z, ()),
    ("weno5 z upward range", "if (weno5 && wk > 0.0 && k >= 3 && k <= nz - 2) {",
     "if (weno5 && wk > 0.0 && k >= 4 && k <= nz - 2) {", ka.check_weno5_z, ()),
    ("weno5 z downward range", "if (weno5 && wk < 0.0 && k >= 2 && k <= nz - 3) {",
     "if (weno5 && wk < 0.0 && k >= 3 && k <= nz - 3) {", ka.check_weno5_z, ()),
    ("weno5 power one", "double a0 = 0.1 / ((eps + b0) * (eps + b0));", "double a0 = 0.1 / (eps + b0);",
     ka.check_weno5_indicators, ()),
]


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    return tmp_path_factory.mktemp("mutants")


@pytest.mark.parametrize("name,old,new,check,args", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pin_catches_planted_mistake(workdir, name, old, new, check, args):
    src = open(SRC).read()
    assert src.count(old) == 1, f"{name}: the text to change must occur exactly once"
    d = workdir / name.replace(" ", "_")
    d.mkdir()
    shutil.copy(HDR, d / "mimadv_oracle.h")
    (d / "mimadv_oracle.c").write_text(src.replace(old, new))
    lib = str(d / "liboracle_mutant.so")
    subprocess.run(["gcc", *CFLAGS, "-o", lib, str(d / "mimadv_oracle.c"), "-lm"], check=True)
    with pytest.raises(AssertionError):
        check(*args, lib=lib)
