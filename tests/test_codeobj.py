"""CPU: register, scratch and memory-instruction use of the library's gfx950 kernels.

Reads the AMDGPU metadata of the code objects built into
libspacecharge_mi355.so (one offload bundle per translation unit in its
.hip_fatbin section) with the ROCm LLVM tools, and checks that no kernel
spills VGPRs and that no register-engine kernel uses scratch memory.
Spills and scratch have cost 20-40% on single kernels (a z convolution held
to three waves per SIMD with the cathode image's extra registers; a
runtime-indexed kernel-argument copy).  The generic mixed-radix engine's
Float64 kernels keep one 16-byte value in scratch in their load loops
(at most 32 bytes are allowed there).
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "spacecharge.jl_amd", "libspacecharge_mi355.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"
# generic-engine Float64 kernels (sc_solve.hip) allowed a small private segment
GENERIC_F64 = re.compile(r"^_ZN2sc12_GLOBAL__N_1\d+k_(rho_fwd_y|zconv|zconv_adj|inv_y|inv_x)Id(E|Li)")
# the Float32 x passes on 512-lane workgroups (x axes beyond 2048 nodes only):
# 256 VGPRs at two waves per SIMD, a few spilled to scratch
WIDE_F32_X = re.compile(r"^_ZN2sc12_GLOBAL__N_1\d+k_(rho_fwd_x|inv_x|green_1)IfLi16ELi512E")


def _tool(name):
    path = os.path.join(LLVM, name)
    return path if os.path.exists(path) else None


def _code_objects(tmp_path):
    """The gfx950 code objects of the library's offload bundles (paths)."""
    objcopy, bundler = (_tool(n) for n in ("llvm-objcopy", "clang-offload-bundler"))
    if not (objcopy and bundler) or not os.path.exists(LIB):
        pytest.skip("ROCm LLVM tools or the built library not available")
    fat = tmp_path / "fat.bin"
    subprocess.run([objcopy, "--dump-section", f".hip_fatbin={fat}", LIB], check=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, "no offload bundle in .hip_fatbin"
    out = []
    for i, s in enumerate(starts):
        e = starts[i + 1] if i + 1 < len(starts) else len(data)
        part = tmp_path / f"c{i}.bin"
        part.write_bytes(data[s:e])
        co = tmp_path / f"c{i}.co"
        subprocess.run([bundler, "--unbundle", "--type=o", f"--input={part}", f"--targets={TARGET}",
                        f"--output={co}"], check=True)
        out.append(co)
    return out


def _kernels(tmp_path):
    objcopy, bundler, readelf = (_tool(n) for n in ("llvm-objcopy", "clang-offload-bundler",
                                                    "llvm-readelf"))
    if not (objcopy and bundler and readelf) or not os.path.exists(LIB):
        pytest.skip("ROCm LLVM tools or the built library not available")
    fat = tmp_path / "fat.bin"
    subprocess.run([objcopy, "--dump-section", f".hip_fatbin={fat}", LIB], check=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, "no offload bundle in .hip_fatbin"
    kernels = {}
    for i, s in enumerate(starts):
        e = starts[i + 1] if i + 1 < len(starts) else len(data)
        part = tmp_path / f"b{i}.bin"
        part.write_bytes(data[s:e])
        co = tmp_path / f"b{i}.co"
        subprocess.run([bundler, "--unbundle", "--type=o", f"--input={part}", f"--targets={TARGET}",
                        f"--output={co}"], check=True)
        notes = subprocess.run([readelf, "--notes", str(co)], check=True, capture_output=True,
                               text=True).stdout
        for block in re.split(r"\n\s+- \.", notes):
            name = re.search(r"(?:^|\n)\s*(?:\.)?name:\s+(_Z\S+)", block)
            if not name or ".kernarg_segment_size" not in block:
                continue

            def field(key):
                m = re.search(r"\." + key + r":\s+(\d+)", block)
                return int(m.group(1)) if m else 0
            kernels[name.group(1)] = {k: field(k) for k in ("vgpr_count", "vgpr_spill_count",
                                                            "sgpr_spill_count",
                                                            "private_segment_fixed_size")}
    return kernels


def test_no_spills_and_no_scratch(tmp_path):
    kernels = _kernels(tmp_path)
    assert len(kernels) > 100, len(kernels)
    assert any("k_zconv_r" in k for k in kernels) and any("k_interp" in k for k in kernels)
    bad = []
    for name, f in kernels.items():
        if WIDE_F32_X.match(name):
            if f["vgpr_spill_count"] > 16 or f["private_segment_fixed_size"] > 64:
                bad.append((name, f))
        elif f["vgpr_spill_count"]:
            bad.append((name, f))
        elif f["private_segment_fixed_size"]:
            if not (GENERIC_F64.match(name) and f["private_segment_fixed_size"] <= 32):
                bad.append((name, f))
    assert not bad, bad


# Kernels that may take their buffers from a descriptor table (grouped
# launches) but run the single-mesh solve too: their loads and stores must be
# global, not flat (flat ones also count against the LDS wait counter, so the
# transforms' LDS exchanges would wait for memory operations in flight).
GLOBAL_ONLY = re.compile(r"k_(rho_fwd_x_r|rho_fwd_y_r|zconv_r|inv_y_r|inv_x_r|green_1r|green_2r|"
                         r"green_3r|green_f3|green_f)I")


def test_transform_and_green_kernels_use_global_memory_instructions(tmp_path):
    objdump = _tool("llvm-objdump")
    if not objdump:
        pytest.skip("llvm-objdump not available")
    seen, bad = 0, []
    for co in _code_objects(tmp_path):
        asm = subprocess.run([objdump, "-d", "--no-show-raw-insn", str(co)], check=True,
                             capture_output=True, text=True).stdout
        for block in re.split(r"\n(?=[0-9a-f]+ <)", asm):
            head = re.match(r"[0-9a-f]+ <(_Z\S+)>:", block)
            if not head or not GLOBAL_ONLY.search(head.group(1)):
                continue
            seen += 1
            n = len(re.findall(r"\bflat_(load|store)", block))
            if n:
                bad.append((head.group(1), n))
    assert seen >= 40, seen
    assert not bad, bad
