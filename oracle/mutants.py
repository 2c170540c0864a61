"""Checks that the oracle's pins bite -- TEST INFRASTRUCTURE ONLY.

Each entry plants one plausible mistake (a shifted band, a cascading read, a
wrong constant or sign, coarser filtering weights, the opposite collision rule)
in a temporary copy of the oracle's sources, builds it, and runs the pins that
cover that function (tests/test_oracle_pins.py, or the test file the entry
names) against the mutated build (QM_ORACLE_LIB). Every mutant must make a pin
fail.

    python oracle/mutants.py [test-file substring]
"""
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MUTS = [
  ("qm_oracle.cpp", "if (sx < 2 || sx > W - 2 || sy < 2 || sy > H - 2) continue;", "if (sx < 1 || sx > W - 1 || sy < 1 || sy > H - 1) continue;", "hole_filling"),
  ("qm_oracle.cpp", "const int32_t v = table[(size_t)sy * W + sx];", "const int32_t v = filled[(size_t)sy * W + sx];", "hole_filling"),
  ("qm_oracle.cpp", "for (int sy = ty - 1; sy <= ty + 1 && filled[t] == 0; ++sy)", "for (int sy = ty + 1; sy >= ty - 1 && filled[t] == 0; --sy)", "hole_filling"),
  ("qm_semidense.cpp", "(did_subpixel ? 0.05 : 0.5)", "(did_subpixel ? 0.5 : 0.05)", "sd_search"),
  ("qm_semidense.cpp", "const float photo = 4.0f * 4.0f / my_gradient / my_gradient;", "const float photo = 3.0f * 3.0f / my_gradient / my_gradient;", "sd_search"),
  ("qm_semidense.cpp", "const float geo = 0.25f * gradient_max * gradient_max / (my_gradient * my_gradient);", "const float geo = 0.25f * gradient_max / my_gradient;", "sd_search"),
  ("qm_semidense.cpp", "      best_match_sub -= d;\n      best_score = (best_score - (2.0f * d)", "      best_match_sub += d;\n      best_score = (best_score - (2.0f * d)", "sd_search"),
  ("qm_semidense.cpp", "const int a = (int)std::rint((xb - fx0) * 256.0f)", "const int a = (int)std::floor((xb - fx0) * 64.0f) * 4", "sd_search or bilinear"),
  ("qm_semidense.cpp", "if (std::fabs(second_best_step - best_step) > 1.0f && kMinDistanceErrorStereo * best_score > second_best_score)", "if (std::fabs(second_best_step - best_step) > 5.0f && kMinDistanceErrorStereo * best_score > second_best_score)", "sd_search"),
  ("qm_semidense.cpp", "            const float w = 1.0f / o.v[5];", "            const float w = 1.0f / (o.v[5] + kRegDistVar);", "sd_fill"),
  ("qm_semidense.cpp", "(s.v[1] >= (float)kMinBlacklist && sum_valid > (float)kValSumMinCreate)", "(s.v[1] >= (float)kMinBlacklist && sum_valid >= (float)kValSumMinCreate)", "sd_fill"),
  ("qm_semidense.cpp", "  if (inc.g2(x, y) < kMinGrad * kMinGrad) continue;\n      Seed s = snap", "  if (inc.g2(x, y) < 0.0f) continue;\n      Seed s = snap", "sd_fill"),
  ("qm_semidense.cpp", "      if (k > t) t = k;", "      if (t == 0 || k < t) t = k;", "sd_propagation"),
  ("qm_semidense.cpp", "  np = se3_mul_point(T, kp);\n  new_idepth = 1.0f / np.z;", "  np = se3_mul_point(T, kp);\n  new_idepth = 1.0f / kp.z;", "sd_propagation"),
  # the upsample's segmented-scan solver, against its FP64 equality with the
  # Thomas order and the dense FP64 solve (tests/test_upsample_scan.py)
  ("qm_oracle.cpp", "    if (k == 0 || k == sg) A[k] = Real(0);\n    R[k] =", "    if (k == 0) A[k] = Real(0);\n    R[k] =", "fp64_scan or dense", "tests/test_upsample_scan.py"),
  ("qm_oracle.cpp", "(row < 2 ? 15 : 47)", "(row < 2 ? 14 : 47)", "fp64_scan or dense", "tests/test_upsample_scan.py"),
  ("qm_oracle.cpp", "return (k == sg - 1 || k == n - 1) ? Real(0) : CH[k];", "return (k == n - 1) ? Real(0) : CH[k];", "fp64_scan or dense", "tests/test_upsample_scan.py"),
  ("qm_oracle.cpp", "(Real)(G[i - 1].g12 / G[i - 1].g22)", "(Real)(G[i - 1].g11 / G[i - 1].g21)", "fp64_scan or dense", "tests/test_upsample_scan.py"),
]
def main():
  res = []
  only = sys.argv[1] if len(sys.argv) > 1 else None  # run only the mutants whose test file contains this
  for i, (f, old, new, k, *tf) in enumerate(MUTS):
    test_file = tf[0] if tf else 'tests/test_oracle_pins.py'
    if only and only not in test_file:
      continue
    tmp = tempfile.mkdtemp(prefix='qm_mutant_')
    for n in ("qm_oracle.cpp", "qm_semidense.cpp", "qm_prior.cpp"):
        shutil.copy(os.path.join(ROOT, 'oracle', n), tmp)
    src = open(os.path.join(tmp, f)).read()
    assert src.count(old) >= 1, (i, old)
    open(os.path.join(tmp, f), 'w').write(src.replace(old, new, 1))
    so = os.path.join(tmp, 'libmut.so')
    subprocess.check_call(['g++', '-O2', '-std=c++17', '-fPIC', '-fopenmp', '-ffp-contract=off', '-fno-fast-math', '-shared', '-o', so,
                           os.path.join(tmp, 'qm_oracle.cpp'), os.path.join(tmp, 'qm_semidense.cpp'), os.path.join(tmp, 'qm_prior.cpp')])
    env = dict(os.environ, QM_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, '-m', 'pytest', test_file, '-q', '-k', k], cwd=ROOT, env=env, capture_output=True, text=True)
    last = r.stdout.strip().split('\n')[-1]
    res.append((i, f, k, 'CAUGHT' if r.returncode != 0 else 'MISSED', last))
    print(res[-1], flush=True)
    shutil.rmtree(tmp, ignore_errors=True)
  missed = [r for r in res if r[3] == 'MISSED']
  print(f'{len(res) - len(missed)} of {len(res)} mutants caught')
  return 1 if missed else 0


if __name__ == '__main__':
    sys.exit(main())
