"""Host-side logic: Tonto's basis conventions, the cluster generator, the SCF driver."""
import math

import numpy as np
import torch

from tonto_amd import basis, geometry
from tonto_amd.scf import run_rhf


def test_component_order_matches_gaussian_data():
    # foofiles/gaussian_data.foo:295-479: Tonto's d and f component orders
    assert basis.cart_components(2) == [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
    assert basis.cart_components(3) == [(3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1),
                                        (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2), (1, 1, 1)]
    for l in range(9):
        c = basis.cart_components(l)
        assert len(c) == len(set(c)) == basis.n_comp(l) and all(sum(x) == l for x in c)


def test_unnormalize_gives_unit_x_power_norm():
    # <x^l e^-ar^2 | x^l e^-br^2> summed with the un-normalised coefficients must be 1
    for name in ("6-31G(d)", "cc-pVTZ"):
        for el in ("O", "H"):
            for l, exps, coefs in basis._library()[name][el]:
                c = basis.unnormalize(l, exps, coefs)
                s = 0.0
                for ai, ci in zip(exps, c):
                    for aj, cj in zip(exps, c):
                        p = ai + aj
                        s += ci * cj * (math.pi / p) ** 1.5 * basis.double_factorial_odd(l) / (2 * p) ** l
                assert abs(s - 1.0) < 1e-12


def test_cluster_sizes_match_the_workloads():
    assert geometry.water_cluster(8, "6-31G(d)").n_bf == 152
    assert geometry.water_cluster(32, "6-311G(d,p)").n_bf == 992
    assert geometry.water_cluster(64, "cc-pVTZ").n_bf == 4160
    m = geometry.water_cluster(20, "6-31G(d)")
    assert m.n_bf == 380


def test_cluster_generator_is_seeded_and_physical():
    s1, x1 = geometry.water_cluster_xyz(8, seed=5)
    s2, x2 = geometry.water_cluster_xyz(8, seed=5)
    _, x3 = geometry.water_cluster_xyz(8, seed=6)
    assert s1 == s2 and np.array_equal(x1, x2) and not np.array_equal(x1, x3)
    for k in range(8):
        o, h1, h2 = x1[3 * k:3 * k + 3]
        assert abs(np.linalg.norm(h1 - o) - 0.96) < 1e-12 and abs(np.linalg.norm(h2 - o) - 0.96) < 1e-12
        ang = math.degrees(math.acos(np.dot(h1 - o, h2 - o) / 0.96 ** 2))
        assert abs(ang - 104.5) < 1e-9


def test_scf_driver_on_a_model_problem():
    # two-level model with a fixed two-electron part: the driver must reach the
    # aufbau solution of h + G(P) self-consistently
    h = torch.tensor([[-1.0, 0.2], [0.2, -0.5]], dtype=torch.float64)
    S = torch.eye(2, dtype=torch.float64)

    def jk(P, mode):
        return 0.3 * P, 0.2 * P
    r = run_rhf(h, S, 2, 0.0, jk, convergence=1e-12, diis_tol=1e-10)
    F = h + 0.3 * r.P - 0.1 * r.P
    e, C = torch.linalg.eigh(F)
    assert r.converged and torch.allclose(r.P, 2 * C[:, :1] @ C[:, :1].T, atol=1e-9)


def test_sixteen_bit_packing_round_trip(tmp_path):
    """Pair records and ket queues pack two 16-bit fields (first basis functions up to
    65535) into one int: the upper field must survive the round trip above 32767, where
    it sets the sign bit.  Host build of the device header's pack16/hi16/lo16."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "pack.cpp"
    src.write_text(
        '#include "tgpu_device.h"\n#include <cstdio>\n'
        "int main() { int bad = 0;\n"
        "  const int v[] = {0, 1, 32767, 32768, 40000, 65535};\n"
        "  for (int hi : v) for (int lo : v) { const int p = pack16(hi, lo);\n"
        "    if (hi16(p) != hi || lo16(p) != lo) ++bad; }\n"
        "  std::printf(\"%d\\n\", bad); return bad != 0; }\n")
    exe = tmp_path / "pack"
    subprocess.check_call(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-std=c++17",
                           "-I", os.path.join(root, "tonto_amd", "csrc"), str(src), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "0"


def test_bench_traffic_table_names_class_pairs():
    """bench.py attaches rocprofv3 traffic to the roofline of the dominant class pair by
    the name it gives that kernel ("jk class (ab|cd) mixed"); every entry of its table must
    name a class pair of the library and carry consistent byte counts."""
    import importlib.util
    import os
    from tonto_amd.flops import class_pairs
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py")
    spec = importlib.util.spec_from_file_location("bench_module", path)
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    names = {"jk class ({}{}|{}{}) mixed".format(*cp) for cp in class_pairs()}
    assert bench.TRAFFIC
    for name, t in bench.TRAFFIC.items():
        assert name in names, name
        assert abs(t["bytes"] - t["read"] - t["written"]) <= 1e-6 * t["bytes"]
        assert t["dispatches"] > 0 and t["read"] > 0 and t["written"] > 0
