"""Every device kernel of the library keeps its private state in registers: no
scratch (private-segment) memory and no register spills, as the gfx950 compiler
reports them (-Rpass-analysis=kernel-resource-usage) with the library's own
flags.  Scratch has cost real HBM traffic here: HIP's float2 is a class type, a
conditional between two float2 lvalues selects their addresses and sends the
register copy to scratch, and a private array indexed at run time lands there
too.  Compiles device code only (CPU, about ten seconds)."""
import concurrent.futures as cf
import os
import re
import subprocess

import pytest

from genfire_amd import build as B

SOURCES = sorted(f for f in os.listdir(B.CSRC) if f.endswith(".hip"))


def _hipcc_or_skip():
    try:
        return B._hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")


def _resources(src):
    cmd = [_hipcc_or_skip(), "--cuda-device-only", "-c", os.path.join(B.CSRC, src), "-o", os.devnull,
           "-Rpass-analysis=kernel-resource-usage"] + B.COMMON + ["-x", "hip"] + B.HIP_FLAGS
    if src in B.NO_CONTRACT:
        cmd += ["-ffp-contract=off"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out, name = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        m = re.search(r"remark:\s+(ScratchSize \[bytes/lane\]|VGPRs Spill|SGPRs Spill): (\d+)", line)
        if m and name:
            out[name][m.group(1)] = int(m.group(2))
    return out


@pytest.fixture(scope="module")
def resources():
    _hipcc_or_skip()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        return dict(zip(SOURCES, ex.map(_resources, SOURCES)))


@pytest.mark.parametrize("src", SOURCES)
def test_kernels_use_no_scratch(resources, src):
    assert sum(len(v) for v in resources.values()) > 50   # the remarks were parsed
    kernels = resources[src]
    if not kernels:
        pytest.skip(f"{src} defines no device kernels")
    bad = {k: v for k, v in kernels.items()
           if v.get("ScratchSize [bytes/lane]", 0) or v.get("VGPRs Spill", 0) or v.get("SGPRs Spill", 0)}
    assert not bad, bad
