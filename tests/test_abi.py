"""C ABI checks that need no GPU: the HIP library loads, exports exactly
what include/octree_pipeline.h declares, rejects bad batch tables on the
host, and fails loudly (no CPU fallback) when no device is present."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import pipeline_amd
from pipeline_amd import clouds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "octree_pipeline.h")


def declared():
    text = open(HEADER).read()
    return set(re.findall(r"^(?:int|const char\*)\s+(ocp_\w+)\s*\(", text, re.M))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T ocp_" in line}


def test_library_exports_every_declared_symbol():
    lib = pipeline_amd.load_library()
    names = declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert exported(pipeline_amd.LIB_PATH) == names
    assert set(pipeline_amd.SIGNATURES) == names


def test_oracle_library_loads():
    L = oracle.lib()
    assert L.orc_max_threads() >= 1


def test_batch_table_checks_on_host():
    ok = pipeline_amd.check_batch([0, 5, 5, 9], [0, 1, 2], [1, 1, 1])
    assert ok == pipeline_amd.OCP_OK
    assert pipeline_amd.check_batch([1, 5], [0], [1]) == 1           # must start at 0
    assert pipeline_amd.check_batch([0, 5, 3], [0, 0], [1, 1]) == 1  # decreasing
    assert pipeline_amd.check_batch([0, 5], [0], [0]) == 1           # zero range
    assert pipeline_amd.check_batch([0, 5], [float("nan")], [1]) == 1
    assert pipeline_amd.check_batch([0, (1 << 27) + 1], [0], [1]) == 3  # segment too large
    assert pipeline_amd.check_batch([0], [], []) == 1               # no segment


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_gpu_fails_loudly():
    with pytest.raises(pipeline_amd.OcpError):
        pipeline_amd.Context(0)


def test_generators_are_seeded_and_clipped():
    a = clouds.uniform(1000, seed=3)
    b = clouds.uniform(1000, seed=3)
    np.testing.assert_array_equal(a, b)
    assert a.shape == (1000, 4) and (a[:, 3] == 0).all()
    c = clouds.clustered(1 << 20, seed=3)
    assert c[:, :3].min() >= 0 and c[:, :3].max() < 1024
    # a duplicate fraction far above a uniform cloud's after 32-bit Morton
    # quantisation (uniform: ~0.05% at this size)
    assert np.unique(oracle.morton(c)).size < 0.97 * c.shape[0]
    pts, offs, mn, rg, kinds = clouds.batch(8, 10, 1000, seed=1)
    assert offs[0] == 0 and offs[-1] == pts.shape[0] and (np.diff(offs) > 0).all()
    for k in range(8):
        seg = pts[offs[k]:offs[k + 1], :3]
        q = (seg - mn[k]) / rg[k] * 1024
        assert q.min() >= 0 and q.max() < 1024
    assert pipeline_amd.check_batch(offs, mn, rg) == 0


def _lcg_points(n=4096, seed=12345):
    """The points tests/c/abi_demo.c generates (32-bit LCG, float32 ops)."""
    out = np.zeros((n, 4), np.float32)
    flat = out.reshape(-1)
    st = seed
    for i in range(n * 4):
        st = (st * 1664525 + 1013904223) & 0xFFFFFFFF
        if i % 4 != 3:
            flat[i] = np.float32(st >> 8) / np.float32(16777216.0) * np.float32(1000.0)
    return out


def test_c_caller_compiles_and_links(tmp_path):
    from pipeline_amd import build as B
    exe = B.build_c_demo(str(tmp_path / "abi_demo"))
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_c_caller_runs():
    from pipeline_amd import build as B
    exp = oracle.build_plain(_lcg_points())
    assert os.path.exists(B.C_DEMO), "build the library first (pipeline_amd/build.py)"
    r = subprocess.run([B.C_DEMO, str(exp["n_unique"]), str(exp["n_oct"])],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def test_fixture_caller_compiles_and_links(tmp_path):
    """include/octree_fixture.hpp compiles with g++ -Wall -Wextra -Werror and
    its caller links against the library."""
    from pipeline_amd import build as B
    exe = B.build_fixture_demo(str(tmp_path / "fixture_demo"))
    assert os.path.exists(exe)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [307_200, 10_000_000])
def test_fixture_matches_fused_build_and_oracle(tmp_path, n):
    """ocp::GpuFixture (benchmarks/cuda/fixture.cuh:24-91 over the C ABI) plus
    the per-stage octree entries give the fused build's bytes (checked inside
    the program), and its counts are the oracle's on the same points."""
    from pipeline_amd import build as B
    assert os.path.exists(B.FIXTURE_DEMO), "build the library first (pipeline_amd/build.py)"
    path = tmp_path / "points.f32"
    r = subprocess.run([B.FIXTURE_DEMO, str(n), str(path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    pts = np.fromfile(path, np.float32).reshape(n, 4)
    assert np.all(pts[:, 3] == 1.0) and pts[:, :3].min() >= 0 and pts[:, :3].max() < 1024
    exp = oracle.build_plain(pts)
    u, k = (int(x) for x in r.stdout.split())
    assert (u, k) == (exp["n_unique"], exp["n_oct"])


def test_workspace_size_query():
    """The workspace query needs no device and grows with the workload."""
    a = pipeline_amd.workspace_size(1_000_000)
    b = pipeline_amd.workspace_size(10_000_000)
    c = pipeline_amd.workspace_size(10_000_000, permutation=True)
    d = pipeline_amd.workspace_size(10_000_000, 512)
    assert 0 < a["build_bytes"] < b["build_bytes"] < c["build_bytes"]
    assert b["build_bytes"] < d["build_bytes"]
    assert a["sort_bytes"] <= b["sort_bytes"] and b["scan_bytes"] > 0
    with pytest.raises(pipeline_amd.OcpError):
        pipeline_amd.workspace_size(-1)


def test_struct_layouts_match_header(tmp_path):
    """The ctypes structures and numpy dtypes of pipeline_amd mirror the
    header's structs: sizes and every field's offset, as a C compiler lays
    them out (a silent mismatch would misread every field after it)."""
    mirrors = {
        "ocp_results": pipeline_amd.Results,
        "ocp_tune": pipeline_amd.Tune,
        "ocp_learned_stats": pipeline_amd.LearnedStats,
        "ocp_workspace_sizes": pipeline_amd.WorkspaceSizes,
        "ocp_tree_report": pipeline_amd.TreeReport,
        "ocp_oct_node": pipeline_amd.OCT_NODE_DTYPE,
        "ocp_compact_node": pipeline_amd.COMPACT_NODE_DTYPE,
    }

    def fields(m):
        if isinstance(m, np.dtype):
            return m.itemsize, {f: m.fields[f][1] for f in m.names}
        return ctypes.sizeof(m), {f: getattr(m, f).offset for f, _ in m._fields_}

    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "octree_pipeline.h"',
             "int main(void) {"]
    for name, m in mirrors.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for f in fields(m)[1]:
            lines.append(f'  printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.dirname(HEADER),
                    str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                       text=True, check=True).stdout.split("\n")
               if line)
    for name, m in mirrors.items():
        size, offs = fields(m)
        assert int(got[name]) == size, name
        for f, off in offs.items():
            assert int(got[f"{name}.{f}"]) == off, (name, f)
