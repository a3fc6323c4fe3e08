"""The oracle's pinning tests catch plausible one-line mistakes in every oracle cell.

Each mutant below is a copy of oracle/ce_oracle.cpp with one mistake planted (a wrong sign, index,
operand, constant or branch), built on the CPU into a temporary directory and loaded by the pinning
tests through CE_ORACLE_LIB in a child process.  The test passes when every mutant makes at least
one of those tests fail.  The committed fixtures are left out of the child run: they were written by
the oracle itself and would catch any change, so they pin nothing.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "ce_oracle.cpp")
PIN_TESTS = ["tests/test_oracle_pins.py", "tests/test_oracle_pipeline.py", "tests/test_oracle_linalg.py",
             "tests/test_oracle_kalman.py", "tests/test_oracle_ref32.py"]

# (cell, text in the oracle, the same text with the mistake)
MUTANTS = [
    ("lina -BU K sign", "lina[(ns + i) * nt + j] = -(P.dbu[i] * K[j]);", "lina[(ns + i) * nt + j] = (P.dbu[i] * K[j]);"),
    ("lina L DC operand", "lina[i * nt + ns + j] = L[i] * P.dc[j];", "lina[i * nt + ns + j] = K[i] * P.dc[j];"),
    ("lina L (C - DK) sign", "lina[i * nt + j] = (P.sa[i * ns + j] - bk) - lcd;",
     "lina[i * nt + j] = (P.sa[i * ns + j] - bk) + lcd;"),
    ("lina C - DK sign", "const double cdk = P.sc[j] - P.sd * K[j];", "const double cdk = P.sc[j] + P.sd * K[j];"),
    ("lina DA index", "lina[(ns + i) * nt + ns + j] = P.da[i * nd + j];", "lina[(ns + i) * nt + ns + j] = P.da[j * nd + i];"),
    ("mode Re scale", "Re[i] = cedm::log(cedm::fmax_(mag, EPS30)) / ts;", "Re[i] = cedm::log(cedm::fmax_(mag, EPS30)) * ts;"),
    ("dedupe conjugates", "cedm::fabs_(cedm::fabs_(Im[i]) - cedm::fabs_(pim)) < EPS6",
     "cedm::fabs_(Im[i] - pim) < EPS6"),
    ("unstable score", "sc[0] = (double)c.unstable_err + cedm::log(max_mag) / ts;\n    sc[1] = sc[0];\n  } else {",
     "sc[0] = (double)c.unstable_err + max_mag / ts;\n    sc[1] = sc[0];\n  } else {"),
    ("yss count", "yss = yss / (double)(T - D.steady_start_id);", "yss = yss / (double)(T - D.steady_start_id + 1);"),
    ("overshoot normalisation", "fr[2] = peak / cedm::fmax_(yss * yss, EPS6) - 1.0;",
     "fr[2] = peak / cedm::fmax_(yss, EPS6) - 1.0;"),
    ("overshoot peak", "peak = cedm::fmax_(peak, y[t] * yss);", "peak = cedm::fmax_(peak, y[t]);"),
    ("wn aggregate", "else { wn = cedm::exp(lw / esum); zeta = zw / esum; }", "else { wn = lw / esum; zeta = zw / esum; }"),
    ("zeta weight", "lw = lw + energy * cedm::log(wn_e);\n        zw = zw + energy * (-re_e / wn_e);",
     "lw = lw + energy * cedm::log(wn_e);\n        zw = zw + (-re_e / wn_e);"),
    ("oscillatory filter", "if (im_e < EPS6) continue;", "if (im_e < -1.0) continue;"),
    ("zeta term", "double beh = lr * lr + ((zeta - dmp) * (zeta - dmp) / dmp) / dmp;",
     "double beh = lr * lr + ((zeta - dmp) * (zeta - dmp) / dmp);"),
    ("target wn", "const double twn = ((2.0 * PI_D) * (double)c.bandwidth)", "const double twn = ((1.0 * PI_D) * (double)c.bandwidth)"),
    ("overshoot hinge", "const double err_os = cedm::fmax_(0.0, fr[2] - os) / os;",
     "const double err_os = cedm::fmax_(0.0, fr[2]) / os;"),
    ("steady-state hinge", "cedm::fabs_(fr[3] / (double)c.ref_amplitude - 1.0) - sse", "cedm::fabs_(fr[3] - 1.0) - sse"),
    ("nominal loop sign", "Acl[i * m + ns + j] = -(L[i] * P.pc[j]);", "Acl[i * m + ns + j] = (L[i] * P.pc[j]);"),
    ("nominal-unstable margin cell", "sc[2] = UE;\n    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;",
     "sc[2] = sc[0];\n    am[0] = am[1] = am[2] = am[3] = am[4] = 0.0;"),
    ("nominal-unstable total", "const double total = cedm::log((sc[0] + UE) + EPS30);",
     "const double total = cedm::log(sc[0] + EPS30);"),
    ("b denominator", "B = cedm::fmin_(B, num / (cedm::sqrt_(1.0 + Ps * Ps) * cedm::sqrt_(1.0 + Cs * Cs)));",
     "B = cedm::fmin_(B, num / cedm::sqrt_(1.0 + Ps * Ps));"),
    ("b plant scaling", "const double Ps = P.alpha * cedm::hypot(P.presp[w].r, P.presp[w].i);",
     "const double Ps = cedm::hypot(P.presp[w].r, P.presp[w].i);"),
    ("PM offset", "double pm_deg = 180.0 + phase_c * RAD2DEG;", "double pm_deg = phase_c * RAD2DEG;"),
    ("DM interpolation", "const double theta_c = prevTheta + frac * (theta_w - prevTheta);",
     "const double theta_c = theta_w;"),
    ("DM units", "DM = cedm::fmin_(DM, ((pm_rad / theta_c) * ts) * 1000.0);", "DM = cedm::fmin_(DM, (pm_rad / theta_c) * 1000.0);"),
    ("GM target side", "const double target = TWO_PI_D * (b > a ? kb : ka);", "const double target = TWO_PI_D * (b > a ? ka : kb);"),
    ("GM interpolation", "const double mag_p = prevMag + frac * (m - prevMag);", "const double mag_p = m;"),
    ("GM sign", "GM = cedm::fmin_(GM, -20.0 * cedm::log10(mag_p));", "GM = cedm::fmin_(GM, 20.0 * cedm::log10(mag_p));"),
    ("sweep failure flag", "am[0] = 0.0; am[1] = 0.0; am[2] = 0.0; am[3] = 1.0; am[4] = 0.0;",
     "am[0] = 0.0; am[1] = 0.0; am[2] = 0.0; am[3] = 0.0; am[4] = 0.0;"),
    ("margin error b term", "const double margin = (err_gm * err_gm + err_pm * err_pm) + err_b * err_b;",
     "const double margin = err_gm * err_gm + err_pm * err_pm;"),
    ("total", "const double total = cedm::log((sc[0] + margin) + EPS30);", "const double total = cedm::log(sc[0] + EPS30);"),
    ("non-finite behaviour", "sc[0] = cedm::isfinite_(beh) ? beh : (double)c.eig_fail_err;", "sc[0] = beh;"),
    ("grid spacing", "const float inc = (cedm::logf_(P.d.freq_max) - cedm::logf_(c.freq_min)) / (float)(F - 1);",
     "const float inc = (cedm::logf_(P.d.freq_max) - cedm::logf_(c.freq_min)) / (float)F;"),
    ("grid in double", "P.z[w] = Cx{(double)co, (double)s};",
     "P.z[w] = Cx{cedm::cos((double)th), cedm::sin((double)th)};"),
    ("sim reference sign", "      const double u_raw = cedm::fma_(N, P.ref[t], -kx);",
     "      const double u_raw = cedm::fma_(N, P.ref[t], kx);"),
    ("kalman gain denominator", "hphr = 1.0 / (hphr + R);", "hphr = 1.0 / hphr;"),
    ("kalman innovation sign", "    dz = -dz;\n", "    dz = dz;\n"),
    ("kalman process noise", "Pn[i * ns + i] = Pn[i * ns + i] + Q[i];", "Pn[i * ns + i] = Pn[i * ns + i];"),
    ("kalman joseph term", "Pm[i * ns + j] = a + (kg[i] * R) * kg[j];", "Pm[i * ns + j] = a;"),
    ("sim observer innovation", "const double yerr = yv - yhat;", "const double yerr = yhat - yv;"),
    ("ref32 TwoSum error term", "const float v = (LO + l) + sl;", "const float v = LO + l;"),
    ("ref32 plant state uncompensated", "for (int i = 0; i < nd; i++) hladd(xp[i], xpl[i], ph[i], pl[i]);",
     "for (int i = 0; i < nd; i++) xp[i] = xp[i] + ph[i];"),
    ("ref32 lina B K sign", "lina[i * nt + j] = ((float)P.sa[i * ns + j] - bk) - lcd;",
     "lina[i * nt + j] = ((float)P.sa[i * ns + j] + bk) - lcd;"),
    ("ref32 lina -BU K sign", "lina[(ns + i) * nt + j] = -((float)P.dbu[i] * K[j]);",
     "lina[(ns + i) * nt + j] = ((float)P.dbu[i] * K[j]);"),
    ("ref32 overshoot normalisation", "fr[2] = peak / fmaxf_(yss * yss, EPS6F) - 1.0f;",
     "fr[2] = peak / fmaxf_(yss, EPS6F) - 1.0f;"),
    ("ref32 energy cross term", "((c0 * c0) * g00 + ((2.0f * c0) * c1) * g01) + (c1 * c1) * g11;",
     "((c0 * c0) * g00 + (c0 * c1) * g01) + (c1 * c1) * g11;"),
    ("ref32 GM sign", "GM = fminf_(GM, -20.0f * log10f_(mag_p));", "GM = fminf_(GM, 20.0f * log10f_(mag_p));"),
    ("ref32 PM offset", "float pm_deg = 180.0f + phase_c * (180.0f / PIF);", "float pm_deg = phase_c * (180.0f / PIF);"),
    ("ref32 b plant scaling", "const float Ps = alpha * hypotf_(pr.r, pr.i)", "const float Ps = hypotf_(pr.r, pr.i)"),
    ("ref32 nominal loop sign", "Acl[i * m + ns + j] = -(L[i] * pc[j]);", "Acl[i * m + ns + j] = (L[i] * pc[j]);"),
    ("ref32 margin total", "const float total = cedm::logf_((sc[0] + margin) + EPS30F);",
     "const float total = cedm::logf_(sc[0] + EPS30F);"),
    ("ref32 plant response", "acc = CxT<float>{acc.r + pc[i] * x[i].r, acc.i + pc[i] * x[i].i};",
     "acc = CxT<float>{acc.r + pc[i] * x[i].r, acc.i - pc[i] * x[i].i};"),
]


# the oracle's own build flags (oracle/build.py) at a lower optimisation level, which changes no
# floating-point result without contraction or fast-math
FLAGS = ["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-mfma", "-mavx2",
         "-fno-math-errno"]


def _build_and_run(job):
    k, name, old, new, tmp = job
    with open(SRC) as f:
        text = f.read()
    assert text.count(old) == 1, "mutant %r: its text occurs %d times in the oracle" % (name, text.count(old))
    text = text.replace(old, new).replace('#include "../', '#include "%s/' % ROOT)
    d = os.path.join(tmp, "m%03d" % k)
    os.makedirs(d, exist_ok=True)
    src = os.path.join(d, "ce_oracle.cpp")
    lib = os.path.join(d, "libce_oracle.so")
    with open(src, "w") as f:
        f.write(text)
    r = subprocess.run(FLAGS + ["-o", lib, src], capture_output=True, text=True)
    if r.returncode != 0:
        return name, None, "build failed: " + r.stderr[-400:]
    env = dict(os.environ, CE_ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider"] + PIN_TESTS,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return name, r.returncode, r.stdout[-300:]


def test_every_mutant_fails_a_pin():
    with tempfile.TemporaryDirectory() as tmp:
        jobs = [(k, n, o, w, tmp) for k, (n, o, w) in enumerate(MUTANTS)]
        workers = max(1, min(8, (os.cpu_count() or 2)))
        with ThreadPoolExecutor(workers) as ex:
            results = list(ex.map(_build_and_run, jobs))
    survived = [(n, out) for n, rc, out in results if rc == 0]
    broken = [(n, out) for n, rc, out in results if rc is None]
    assert not broken, broken
    assert not survived, "mutants not caught by any pin: %s" % [n for n, _ in survived]
    # exit code 1 = tests failed (not an import or collection error, which would pin nothing)
    odd = [(n, rc, out) for n, rc, out in results if rc != 1]
    assert not odd, odd
