"""The drop-in boundary from the reference's side: its GPU driver, mcb/mcb_cycle_deg_two_remove.cpp, compiles
unchanged with g++ against include/reference_shim (which stands in for the reference's gpu/common.cuh and
gpu/gpu_struct.cuh) and links against libmcb_hip.so and the reference's own host objects, with every device
symbol it needs resolved by the library.  Runs in a container that has the reference tree (it is read, compiled
into a temporary directory, never run); skipped elsewhere."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
LIB = os.path.join(ROOT, "minimumcyclebasis_amd", "lib")
DRIVER = os.path.join(REF, "mcb", "mcb_cycle_deg_two_remove.cpp")
# the host sources the driver links (mcb/Makefile builds the same set); compressed_trees.cpp at -O0 because its
# accessors lack a return statement (undefined behaviour that -O2 turns into a crash)
HOST = ("mmio", "Host_Timer", "CsrGraphSpanningTree", "CsrGraphDegreeTwoRemoval", "CsrGraphMultiSpanningTree", "FVS")
FLAGS = ["-O2", "-std=c++11", "-fopenmp", "-DMULTI_THREAD", "-DNDEBUG", "-w", "-fpermissive", "-include", "cmath"]
# the device entry points the driver reaches: the five of gpu/common.cuh:18-27 and gpu_struct's through the shim
NEEDED = ("init_cuda", "configure_grid", "allocate_pinned_memory", "free_pinned_memory", "calculate_chunk_size",
          "mcb_gpu_create", "mcb_gpu_initialize_memory", "mcb_gpu_copy_support_vector",
          "mcb_gpu_process_shortest_path", "mcb_gpu_destroy", "mcb_hip_last_error")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER) or shutil.which("g++") is None,
                                reason="needs the reference tree and g++")


def _cc(args):
    r = subprocess.run(["g++", *args], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_reference_driver_builds_against_the_shim(tmp_path):
    if not os.path.exists(os.path.join(LIB, "libmcb_hip.so")):
        pytest.skip("libmcb_hip.so is not built")
    inc = ["-I" + os.path.join(ROOT, "include", "reference_shim"), "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(REF, "include"), "-I" + REF]
    jobs = [FLAGS + inc + ["-c", DRIVER, "-o", str(tmp_path / "driver.o")]]
    jobs += [FLAGS + inc + ["-c", os.path.join(REF, "include", s + ".cpp"), "-o", str(tmp_path / (s + ".o"))]
             for s in HOST]
    jobs.append(FLAGS + ["-O0"] + inc + ["-c", os.path.join(REF, "include", "compressed_trees.cpp"), "-o",
                                         str(tmp_path / "compressed_trees.o")])
    with ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(_cc, jobs))
    exe = tmp_path / "mcb_hip_driver"
    objs = [str(tmp_path / "driver.o")] + [str(tmp_path / (s + ".o")) for s in HOST + ("compressed_trees",)]
    _cc(["-fopenmp", "-o", str(exe), *objs, "-L" + LIB, "-lmcb_hip", "-Wl,-rpath," + LIB,
         "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,--no-undefined"])
    undefined = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True, check=True).stdout
    listed = {line.split()[-1] for line in undefined.splitlines() if line.strip()}
    exported = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIB, "libmcb_hip.so")],
                              capture_output=True, text=True, check=True).stdout
    have = {line.split()[-1] for line in exported.splitlines() if line.strip()}
    for sym in NEEDED:
        assert sym in listed, f"the driver does not call {sym}"
        assert sym in have, f"libmcb_hip.so does not export {sym}"
