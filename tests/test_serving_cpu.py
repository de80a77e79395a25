"""Host-side checks of the serving API: struct layouts of the C header against the ctypes bindings,
the workload generators, and error returns that need no GPU."""
import ctypes
import importlib
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def _serving():
    return importlib.import_module("halide-sar-app_amd.serving")


def _c_layout(tmp_path, structs):
    """sizeof and field offsets of C structs, from a program built against include/sarbp_hip.h."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sarbp_hip.h"', 'int main(void) {']
    for name, fields in structs.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


def test_struct_layouts_match_header(tmp_path, sar):
    sv = _serving()
    pairs = {"sarbp_server_config": sv.ServerConfig, "sarbp_job": sv.JobDesc,
             "sarbp_job_stats": sv.JobStats, "sarbp_plan_desc": sar.PlanDesc}
    lay = _c_layout(tmp_path, {c: [f[0] for f in py._fields_] for c, py in pairs.items()})
    for cname, py in pairs.items():
        assert lay[cname] == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert lay[f"{cname}.{f[0]}"] == getattr(py, f[0]).offset, (cname, f[0])


def test_c4_specs_cover_the_stated_ranges():
    sp = _serving().c4_specs()
    assert len(sp) == 512
    assert {s["ns"] for s in sp} == {424, 1024, 2048}
    assert {s["side"] for s in sp} == {256, 512, 1024, 2048}
    npl = np.array([s["npulses"] for s in sp])
    tl = np.array([s["taylor"] for s in sp])
    assert npl.min() >= 200 and npl.max() <= 4000 and tl.min() >= 17 and tl.max() <= 35
    assert _serving().c4_specs() == sp  # seeded


def test_c5_specs_interleave_large_jobs():
    sp = _serving().c5_specs()
    big = [i for i, s in enumerate(sp) if s.get("big")]
    assert len(sp) == 2004 and len(big) == 4
    assert all(sp[i]["side"] == 8192 and sp[i]["npulses"] == 6000 for i in big)
    assert 0 < big[0] and big[-1] < len(sp) - 1 and min(np.diff(big)) > 100


def test_server_without_gpu_fails_cleanly(sar):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    sv = _serving()
    with pytest.raises(sv.ServingError) as e:
        sv.Server()
    assert e.value.code == sar.ERR["no_device_interface"]


def test_submit_rejects_null_server_and_job(sar):
    sv = _serving()
    lib = sar.lib()
    jid = ctypes.c_uint64()
    assert lib.sarbp_server_submit(None, None, ctypes.byref(jid)) == sar.ERR["buffer_argument_is_null"]
    st = sv.JobStats()
    assert lib.sarbp_server_wait(None, 1, 0, ctypes.byref(st)) == sar.ERR["buffer_argument_is_null"]
    assert lib.sarbp_server_destroy(None) == 0
    cfg = sv.ServerConfig()
    lib.sarbp_server_config_init(ctypes.byref(cfg))
    assert (cfg.max_inflight, cfg.num_streams, cfg.autotune) == (64, 2, 1)
    assert cfg.hbm_budget_bytes == 64 << 30
    assert cfg.band_order == 0  # least work left first
    cfg.num_streams = 0
    h = ctypes.c_void_p()
    assert lib.sarbp_server_create(ctypes.byref(cfg), ctypes.byref(h)) == sar.ERR["constraint_violated"]
    cfg.num_streams, cfg.band_order = 2, 2
    assert lib.sarbp_server_create(ctypes.byref(cfg), ctypes.byref(h)) == sar.ERR["constraint_violated"]
