"""BVH topology on the CPU: the per-node routines of k_build_chunks (in-chunk
nodes by binary lifting over the chunk's adjacent common-prefix table) and
k_top_boxes (nodes whose range leaves the chunk, by 16-way searches over all
keys), csrc/gi_karras.h, host and device code, compiled for the host and run
over every node of Morton, uniform, duplicate-heavy, clustered, tiny and
batch (mesh id above the Morton code, 32 and 64 bits) key sets, at chunk
sizes 64/256/512.
Each node must equal a plain Karras build (Karras 2012), and the tree must be
well formed (every node reached once, children splitting their parent's
range, ropes = pre-order successors). The reference's BVH (cuBQL's builder
with its level-cut extension, include/third-party/cubql/level_cut_builder.h:
16-364) is internal state whose shape never reaches the output, so the tree
is checked against Karras' definition rather than against it. Planting an
off-by-one in the lifted range end or in the split position fails this
test."""
import os
import subprocess

from gi_test_util import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "karras_test.cpp")
CSRC = os.path.join(ROOT, "gpu-cubql-intersector-standalone_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_chunk_node_routine_equals_plain_karras_and_tree_is_well_formed(tmp_path):
    exe = str(tmp_path / "karras_test")
    build = subprocess.run([HIPCC, "-std=c++17", "-O2", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                            "-I" + CSRC, SRC, "-o", exe], capture_output=True, text=True)
    assert build.returncode == 0, build.stdout + build.stderr
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0 and run.stdout.rstrip().endswith("ALL OK"), run.stdout[-3000:] + run.stderr
    # every key set ran at every chunk size
    for name in ("sphere_morton", "uniform", "duplicates", "clustered", "small", "forest32", "forest64"):
        for chunk in (64, 256, 512):
            assert f"ok {name} " in run.stdout and f"chunk={chunk} " in run.stdout
