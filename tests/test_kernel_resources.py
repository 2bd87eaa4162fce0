"""Register and scratch budgets of the built kernels, read from the gfx950 code object inside libmosegpu.so (its
AMDGPU metadata note: .vgpr_count, .vgpr_spill_count, .private_segment_fixed_size).

The stage kernel's speed rests on an occupancy cliff: at <= 128 VGPRs two 512-thread workgroups of an 8 x 8 x 4
tile share a CU (4 waves per SIMD); at 129 only one fits (2 waves per SIMD), and a stage takes about 16 % longer.
These tests fail the build when a change crosses it, before any GPU run.
"""
import os
import re
import shutil
import subprocess

import pytest

import mosegpu

LLVM = "/opt/rocm/llvm/bin"


def _tool(name):
    p = os.path.join(LLVM, name)
    return p if os.path.exists(p) else shutil.which(name)


@pytest.fixture(scope="module")
def kernels(tmp_path_factory):
    lib = str(mosegpu.LIB_PATH)
    tools = [_tool(t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf")]
    if not os.path.exists(lib) or None in tools:
        pytest.skip("built library or LLVM tools missing")
    objcopy, bundler, readelf = tools
    d = tmp_path_factory.mktemp("co")
    fat = str(d / "fat.bin")
    subprocess.run([objcopy, "--dump-section=.hip_fatbin=" + fat, lib, str(d / "host.o")], check=True,
                   capture_output=True)
    # one offload bundle per translation unit with device code, concatenated by the linker
    data = open(fat, "rb").read()
    starts = [m.start() for m in re.finditer(b"__CLANG_OFFLOAD_BUNDLE__", data)] + [len(data)]
    out = {}
    for n, (a, b) in enumerate(zip(starts[:-1], starts[1:])):
        part, co = str(d / f"fat{n}.bin"), str(d / f"gfx950_{n}.co")
        with open(part, "wb") as f:
            f.write(data[a:b])
        subprocess.run([bundler, "--unbundle", "--type=o", "--input=" + part,
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + co], check=True,
                       capture_output=True)
        notes = subprocess.run([readelf, "--notes", co], check=True, capture_output=True, text=True).stdout
        for item in re.split(r"\n  - \.", notes)[1:]:
            kv = {}
            for line in ("." + item).splitlines():
                m = re.match(r"^(?:    )?\.([a-z_]+):\s+(\S+)", line) if not line.startswith("      ") else None
                if m:
                    kv.setdefault(m.group(1), m.group(2))
            if "name" in kv:
                out[kv["name"]] = kv
    assert out, "no kernel metadata found"
    return out


def _stage(kernels, nr, visc, rs, sh, fast=False):
    kern = "12k_stage_fast" if fast else "7k_stage"
    name = f"_ZN4mose{kern}ILi1ELi{nr}ELb{int(visc)}ELi{rs}ELi512ELi{sh}EEEvNS_7CaseDevENS_9StageArgsE"
    assert name in kernels, name
    return kernels[name]


@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("visc", [True, False])
@pytest.mark.parametrize("rs", [1, 4, 12])
def test_stage_8x8x4_fits_two_workgroups_per_cu(kernels, visc, rs, fast):
    k = _stage(kernels, 0, visc, rs, 1, fast)
    assert int(k["vgpr_count"]) <= 128
    assert int(k["vgpr_spill_count"]) == 0 and int(k["private_segment_fixed_size"]) == 0


@pytest.mark.parametrize("nr", [0, 1])
@pytest.mark.parametrize("sh", [0, 1, 2])
def test_stage_kernels_do_not_spill(kernels, nr, sh):
    for visc in (True, False):
        for rs in (1, 4, 12):
            for fast in (False, True):
                k = _stage(kernels, nr, visc, rs, sh, fast)
                assert int(k["vgpr_spill_count"]) == 0, (nr, sh, visc, rs, fast)


def test_chemistry_occupancy(kernels):
    """k_chem is built for 4 waves per SIMD (its spills were measured to cost less than halving the waves);
    k_radau holds three LU rows and nine vectors per lane and runs at one wave per SIMD."""
    chem = kernels["_ZN4mose6k_chemILi9EEEvNS_7ChemDevENS_8ChemArgsE"]
    assert int(chem["vgpr_count"]) <= 128
    radau = kernels["_ZN4mose7k_radauILi9EEEvNS_7ChemDevENS_8ChemArgsE"]
    assert int(radau["vgpr_count"]) <= 512
