"""Build the native libraries in-tree.

  hicops_amd/libhcx.so       HIP kernels + C ABI (include/hcx.h), gfx950
  hicops_amd/libhcx_dslim.so the reference's C++ entry points over the C ABI
                             (include/hcx_dslim.hpp)
  hicops_amd/libhcx_host.so  synthetic data generator (include/hcx_datagen.h) and the MS2
                             reader (include/hcx_ms2.h)
  oracle/liboracle.so        CPU restatement used by tests and the CPU baseline
  tests/cpp/dslim_driver     test program driving include/hcx_dslim.hpp
  tests/cpp/tsv_driver       the same, leaving PSMs to the library's TSV writer

Run as ``python -m hicops_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"

HIP_SOURCES = ["hcx_api.hip", "k_index.hip", "k_query.hip", "fit_host.hip", "k_chunk.hip", "k_packed.hip", "k_mlp.hip", "k_pick.hip", "k_lbe.hip"]
ARCH = os.environ.get("HCX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every multiply and add rounded separately: fragment bins, hyperscores
    # and the tail-fit regression must match the host's IEEE arithmetic
    "-ffp-contract=off",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def build_host() -> Path:
    out = PKG / "libhcx_host.so"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
          "-o", str(out), str(CSRC / "datagen.cpp"), str(CSRC / "ms2.cpp")])
    return out


def build_hip() -> Path:
    out = PKG / "libhcx.so"
    srcs = [str(CSRC / s) for s in HIP_SOURCES if (CSRC / s).exists()]
    tmp = out.with_suffix(".so.tmp")
    _run(["hipcc", *HIPCC_FLAGS, "-o", str(tmp), *srcs])
    os.replace(tmp, out)
    return out


def build_mirror() -> Path:
    out = PKG / "libhcx_dslim.so"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-I", str(ROOT / "include"),
          "-o", str(out), str(PKG / "host" / "dslim_hcx.cpp"),
          "-L", str(PKG), "-lhcx", "-Wl,-rpath,$ORIGIN"])
    drv = ROOT / "tests" / "cpp" / "dslim_driver"
    # -rdynamic: the driver's DFile_PrintScore takes precedence over the library's
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-rdynamic", "-I", str(ROOT / "include"),
          "-o", str(drv), str(ROOT / "tests" / "cpp" / "dslim_driver.cpp"),
          "-L", str(PKG), "-lhcx_dslim", "-lhcx", "-Wl,-rpath,$ORIGIN/../../hicops_amd"])
    # the same driver leaving PSMs to the library's TSV writer
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-DHCX_TSV_DRIVER", "-I", str(ROOT / "include"),
          "-o", str(drv.with_name("tsv_driver")), str(ROOT / "tests" / "cpp" / "dslim_driver.cpp"),
          "-L", str(PKG), "-lhcx_dslim", "-lhcx", "-Wl,-rpath,$ORIGIN/../../hicops_amd"])
    return out


def build_all() -> None:
    build_oracle()
    build_host()
    build_hip()
    build_mirror()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "oracle": build_oracle, "host": build_host, "hip": build_hip,
         "mirror": build_mirror}[w]()
