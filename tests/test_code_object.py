"""Kernel resources of the built library, read from its gfx950 code objects
(no device needed): the device guess kernels keep everything in registers,
LDS and the rows they read (no private segment and no dynamic stack: no stack frames, no
out-of-line calls whose arguments live in scratch), and the production LM kernel of the metric
workload (pseudo-Voigt, 8 lanes, streamed, compensated, FP64 solve, guessed
starts) neither spills nor uses scratch."""
import os
import re
import shutil
import subprocess

import pytest

LLVM = "/opt/rocm/llvm/bin"
LIB = os.path.join(os.path.dirname(__file__), "..", "batchfit_amd", "libbatchfit_hip.so")
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def kernel_resources(tmp_path):
    """{kernel symbol: {"private": bytes, "spill": VGPRs}} over every bundle of
    the library's .hip_fatbin section (one bundle per translation unit)."""
    fat = tmp_path / "fat.bin"
    subprocess.run([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={fat}", LIB,
                    str(tmp_path / "rest.so")], check=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    kernels = {}
    for i, a in enumerate(starts):
        b = starts[i + 1] if i + 1 < len(starts) else len(data)
        part, co = tmp_path / f"b{i}.bin", tmp_path / f"b{i}.co"
        part.write_bytes(data[a:b])
        subprocess.run([f"{LLVM}/clang-offload-bundler", "--unbundle", "--type=o",
                        f"--targets={TARGET}", f"--input={part}", f"--output={co}"], check=True)
        notes = subprocess.run([f"{LLVM}/llvm-readelf", "--notes", str(co)], check=True,
                               capture_output=True, text=True).stdout
        name = None
        for line in notes.splitlines():
            line = line.strip()
            if line.startswith(".name:"):
                name = line.split(":", 1)[1].strip()
                kernels[name] = {}
            elif name and line.startswith(".private_segment_fixed_size:"):
                kernels[name]["private"] = int(line.split(":")[1])
            elif name and line.startswith(".vgpr_spill_count:"):
                kernels[name]["spill"] = int(line.split(":")[1])
            elif name and line.startswith(".uses_dynamic_stack:"):
                kernels[name]["dynamic_stack"] = line.split(":")[1].strip() == "true"
    return kernels


@pytest.fixture(scope="module")
def resources(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    for tool in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf"):
        if not shutil.which(f"{LLVM}/{tool}"):
            pytest.skip(f"{tool} not found")
    return kernel_resources(tmp_path_factory.mktemp("code_object"))


def test_guess_kernels_use_no_private_memory(resources):
    guess = {k: v for k, v in resources.items() if "guess_kernel" in k or "guess_ragged_kernel" in k}
    assert len(guess) == 8, sorted(guess)  # uniform and ragged, four models
    # The guess of the single-peak models: one kernel per model for edge
    # windows up to 32, 64 and 128 values (sorting networks, then the peak),
    # and beyond, edge levels (bisection in 64-lane groups up to 1024, one
    # wave per fit with the keys read from memory beyond, one kernel per
    # parameter count) then the peak.
    staged = {k: v for k, v in resources.items()
              if "guess_levels_kernel" in k or "guess_levels_stream_kernel" in k
              or "guess_single_kernel" in k or "guess_peak_kernel" in k}
    assert len(staged) == 16, sorted(staged)
    # The ragged form of the one-launch kernel (single-peak fits up to
    # 1,289 points of a ragged batch).
    ragged = {k: v for k, v in resources.items() if "guess_ragged_single_kernel" in k}
    assert len(ragged) == 9, sorted(ragged)
    staged.update(ragged)
    # The wide guess (spectra beyond the one-wave guess's LDS), uniform and
    # ragged, four models.
    wide = {k: v for k, v in resources.items()
            if "guess_wide_kernel" in k or "guess_ragged_wide_kernel" in k}
    assert len(wide) == 8, sorted(wide)
    staged.update(wide)
    guess.update(staged)
    for name, r in guess.items():
        assert r.get("private", 0) == 0 and r.get("spill", 0) == 0, (name, r)
        assert not r.get("dynamic_stack", False), (name, r)


def test_metric_lm_kernel_has_no_spills(resources):
    name = "_ZN5bfhip14lm_pool_kernelILi2ELi8ELb1ELb1ELb0ELb0EEEvNS_7FitArgsE"
    assert name in resources
    assert resources[name].get("private", 0) == 0 and resources[name].get("spill", 0) == 0


def test_fixed_baseline_two_peak_kernels_do_not_spill(resources):
    """The two-peak pool kernels that sum only the six free parameters' rows
    (baseline fixed) keep their state in registers: no private segment."""
    fixb = {k: v for k, v in resources.items()
            if (k.startswith("_ZN5bfhip14lm_pool_kernelILi3E") and k.endswith("Lb1EEEvNS_7FitArgsE"))
            or (k.startswith("_ZN5bfhip21lm_ragged_pool_kernelILi3E")
                and k.endswith("Lb1EEEvNS_10RaggedArgsE"))}
    assert len(fixb) >= 10, sorted(fixb)
    for name, r in fixb.items():
        assert r.get("private", 0) == 0 and r.get("spill", 0) == 0, (name, r)
