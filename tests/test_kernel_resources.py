"""Register, spill and LDS budgets of the built gfx950 kernels (CPU only).

The fused outer-iteration kernels with 8-wave tiles rely on two workgroups
sharing a CU (4 waves per SIMD, at most 128 VGPRs each): one workgroup's
barriers are covered by the other's arithmetic.  A kernel that grows past 128
VGPRs still runs correctly but at one workgroup per CU (a 2048x2048x20 GAP
solve with the per-iteration PSNR went from 32 to 48 ms that way), so the
budgets are checked here, from the metadata of the code objects in the built
library, with ROCm's own LLVM tools (clang-offload-bundler, llvm-readelf).
"""
import os
import re
import shutil
import subprocess

import pytest

from pnp_sci import _lib

LLVM = "/opt/rocm/llvm/bin"
BUNDLER = os.path.join(LLVM, "clang-offload-bundler")
READELF = os.path.join(LLVM, "llvm-readelf")
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def _kernels(tmp_path):
    """{kernel symbol: metadata dict} over every gfx950 code object of the library."""
    lib = _lib.LIB_PATH
    if not os.path.exists(lib):
        pytest.skip("libscigap.so is not built")
    for tool in (BUNDLER, READELF):
        if not os.path.exists(tool):
            pytest.skip("%s is not installed" % tool)
    if shutil.which("objcopy") is None:
        pytest.skip("objcopy is not installed")
    # objcopy rewrites the file it reads: give it a copy, never the library
    # that this process (and the other tests) have mapped.
    copy = tmp_path / "libscigap.so"
    shutil.copyfile(lib, copy)
    fat = tmp_path / "fatbin.bin"
    subprocess.run(["objcopy", "--dump-section", ".hip_fatbin=%s" % fat, str(copy)], check=True,
                   capture_output=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, "no offload bundle in .hip_fatbin"
    kernels = {}
    for i, s in enumerate(starts):
        part = tmp_path / ("bundle%d.bin" % i)
        part.write_bytes(data[s:starts[i + 1] if i + 1 < len(starts) else len(data)])
        co = tmp_path / ("co%d.o" % i)
        r = subprocess.run([BUNDLER, "--type=o", "--input=%s" % part, "--targets=%s" % TARGET,
                            "--output=%s" % co, "--unbundle"], capture_output=True)
        if r.returncode != 0 or not co.exists() or co.stat().st_size == 0:
            continue
        notes = subprocess.run([READELF, "--notes", str(co)], check=True, capture_output=True,
                               text=True).stdout
        # amdhsa.kernels: one YAML mapping per kernel, each starting with "  - ." at
        # an indent of two (argument lists are indented further)
        for block in re.split(r"\n  - \.", notes)[1:]:
            fields = dict(re.findall(r"\.(\w+):\s+(\S+)", "." + block))
            if "symbol" in fields and fields["symbol"].endswith(".kd"):
                kernels[fields["symbol"][:-3]] = fields
    assert kernels, "no kernel metadata found"
    return kernels


def _pick(kernels, pattern):
    rx = re.compile(pattern)
    return {k: v for k, v in kernels.items() if rx.search(k)}


def test_two_workgroup_kernels_fit_128_vgprs_without_spills(tmp_path):
    kernels = _kernels(tmp_path)
    # k_fused<D<=4, BY=8, ...> (all modes, with and without the per-iteration
    # error, both engines), the repair k_fused_fix<D<=4, 8, ...>, and the
    # batched k_fused_batch<4|5, 8, ...>.
    groups = {
        "k_fused<D<=4,8>": _pick(kernels, r"7k_fusedILi[1-4]ELi8ELi1E"),
        "k_fused_fix<D<=4,8>": _pick(kernels, r"11k_fused_fixILi[1-4]ELi8ELi1E"),
        "k_fused_batch<4|5,8>": _pick(kernels, r"13k_fused_batchILi[45]ELi8ELi1E"),
    }
    for name, group in groups.items():
        assert group, "no %s kernels in the library" % name
        for sym, f in group.items():
            assert int(f["vgpr_count"]) <= 128, (name, sym, f["vgpr_count"])
            assert int(f["vgpr_spill_count"]) == 0, (name, sym, f["vgpr_spill_count"])
            # two workgroups per CU also need their LDS: 160 KB per CU
            assert int(f["group_segment_fixed_size"]) <= 80 * 1024, (name, sym)


def test_hot_kernels_do_not_spill(tmp_path):
    kernels = _kernels(tmp_path)
    hot = _pick(kernels, r"k_fused|k_tv_bwd|k_tv_chunk|k_decide")
    assert len(hot) > 50
    for sym, f in hot.items():
        assert int(f["vgpr_spill_count"]) == 0, (sym, f["vgpr_spill_count"])
        assert int(f.get("private_segment_fixed_size", 0)) == 0, sym
    bwd = _pick(kernels, r"k_tv_bwd")
    assert bwd
    for sym, f in bwd.items():  # one 16-wave workgroup per CU: LDS within 160 KB
        assert int(f["group_segment_fixed_size"]) <= 160 * 1024, sym
