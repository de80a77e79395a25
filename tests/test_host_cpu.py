"""Host side of the `sarbp` driver, on the CPU: the platform-directory loader against numpy's own
reading of the same files (synthetic directories of every accepted dtype, the reference's example
data when present), its error cases, the PNG writer decoded independently, and the driver's command
line (everything before the first GPU call)."""
import ctypes
import importlib
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from conftest import ROOT

REF_DATA = "/root/reference/data"


def _platform():
    return importlib.import_module("halide-sar-app_amd.platform_data")


def _write(d, **arrays):
    os.makedirs(d, exist_ok=True)
    for k, v in arrays.items():
        np.save(os.path.join(d, k + ".npy"), v)


def _minimal(d, ns=24, np_=10, seed=0, **over):
    rng = np.random.default_rng(seed)
    a = dict(delta_r=np.array([0.25]), nsamples=np.array([ns], np.int64),
             npulses=np.array([np_], np.int64),
             k_r=rng.uniform(380, 400, ns).astype(np.float32),
             R_c=np.array([1000.0, 20.0, 500.0], np.float32),
             pos=rng.normal(size=(np_, 3)).astype(np.float32),
             phs=(rng.normal(size=(np_, ns)) + 1j * rng.normal(size=(np_, ns))).astype(np.complex64))
    a.update(over)
    a = {k: v for k, v in a.items() if v is not None}
    _write(d, **a)
    return a


def test_loader_reads_required_fields_exactly(tmp_path):
    a = _minimal(tmp_path)
    p = _platform().load(str(tmp_path))
    assert (p["nsamples"], p["npulses"], p["delta_r"]) == (24, 10, 0.25)
    for k in ("k_r", "R_c", "pos", "phs"):
        assert p[k].dtype == a[k].dtype and np.array_equal(p[k], a[k]), k
    for k in ("B", "B_IF", "delta_t", "chirprate", "f_0", "vp", "freq", "k_y", "n_hat", "t"):
        assert p[k] is None, k


def test_loader_downcasts_like_the_reference(tmp_path, capfd):
    """f64 k_r/R_c/pos/freq/B_IF and complex128 phs are rounded to f32 (PlatformData.cpp:38-41,
    120-126, 140-146, 187-193, 222-228, 241-248); int32 counts are read as they are."""
    rng = np.random.default_rng(3)
    ns, np_ = 17, 7
    a = _minimal(tmp_path, ns, np_, nsamples=np.array([ns], np.int32),
                 npulses=np.array([np_], np.int32),
                 k_r=rng.uniform(380, 400, ns), R_c=np.array([1e3, 2.5, 7.25]),
                 pos=rng.normal(size=(np_, 3)) * 1e4,
                 phs=rng.normal(size=(np_, ns)) + 1j * rng.normal(size=(np_, ns)),
                 freq=rng.uniform(9e9, 1e10, ns), B_IF=np.array([6.2e8]),
                 B=np.array([1e9]), delta_t=np.array([1e-3]), chirprate=np.array([9e14]),
                 f_0=np.array([9.6e9]), vp=np.array([120.0]),
                 k_y=rng.normal(size=np_), n_hat=np.array([0, 0, 1], np.float32),
                 t=rng.normal(size=ns))
    p = _platform().load(str(tmp_path))
    for k in ("k_r", "R_c", "pos", "freq"):
        assert p[k].dtype == np.float32 and np.array_equal(p[k], a[k].astype(np.float32)), k
    assert np.array_equal(p["phs"], a["phs"].astype(np.complex64))
    assert p["B_IF"] == np.float32(6.2e8)
    assert (p["B"], p["delta_t"], p["chirprate"], p["f_0"], p["vp"]) == (1e9, 1e-3, 9e14, 9.6e9, 120.0)
    assert np.array_equal(p["k_y"], a["k_y"]) and np.array_equal(p["t"], a["t"])
    assert np.array_equal(p["n_hat"], a["n_hat"])
    out = capfd.readouterr().out
    for k in ("B_IF", "freq", "k_r", "R_c", "pos"):
        assert f"PlatformData: downcasting {k} from double to float" in out
    assert "downcasting phs from complex<double> to complex<float>" in out


@pytest.mark.parametrize("case,message", [
    (dict(pos=np.zeros((10, 2), np.float32)), "Bad shape: pos"),
    (dict(phs=np.zeros((24, 10), np.complex64)), "Bad shape: phs"),
    (dict(k_r=np.zeros(23, np.float32)), "Bad shape: k_r"),
    (dict(R_c=np.zeros(4, np.float32)), "Bad shape: R_c"),
    (dict(delta_r=np.array([0.25], np.float32)), "Bad word size: delta_r"),
    (dict(nsamples=np.array([24], np.int16)), "Bad word size: nsamples"),
    (dict(k_r=np.zeros(24, np.float16)), "Bad word size: k_r"),
    (dict(phs=np.zeros((10, 24), np.float64)), "Bad element type: phs"),
    (dict(n_hat=np.zeros(3, np.float64)), "Bad word size: n_hat"),
    (dict(k_y=np.zeros(9)), "Bad shape: k_y"),
    (dict(t=np.zeros(24, np.float32)), "Bad word size: t"),
])
def test_loader_rejects_bad_files(tmp_path, case, message):
    _minimal(tmp_path, **case)
    with pytest.raises(_platform().PlatformError) as e:
        _platform().load(str(tmp_path))
    assert message in str(e.value)


def test_loader_rejects_missing_big_endian_and_fortran(tmp_path):
    pl = _platform()
    _minimal(tmp_path, phs=None)
    with pytest.raises(pl.PlatformError, match="phs.npy: cannot open"):
        pl.load(str(tmp_path))
    _minimal(tmp_path, k_r=np.zeros(24, ">f4"))
    with pytest.raises(pl.PlatformError, match="big-endian"):
        pl.load(str(tmp_path))
    _minimal(tmp_path, pos=np.asfortranarray(np.zeros((10, 3), np.float32)))
    with pytest.raises(pl.PlatformError, match="Fortran"):
        pl.load(str(tmp_path))


def test_loader_reads_npy_format_2(tmp_path):
    """Format 2.0 has a 4-byte header length (numpy writes it for headers over 64 KiB)."""
    a = _minimal(tmp_path)
    with open(os.path.join(tmp_path, "k_r.npy"), "wb") as f:
        np.lib.format.write_array(f, a["k_r"], version=(2, 0))
    assert np.array_equal(_platform().load(str(tmp_path))["k_r"], a["k_r"])


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's data directory is absent")
def test_loader_matches_numpy_on_reference_data(tmp_path):
    """The reference's example platform directories (AFRL pass 1 in full; Sandia, whose phase
    history is not in the tree, with a stand-in phs of its shape), read in place."""
    pl = _platform()
    afrl = os.path.join(REF_DATA, "AFRL/pass1/HH_npy")
    p = pl.load(afrl)
    ref = {f[:-4]: np.load(os.path.join(afrl, f)) for f in os.listdir(afrl)}
    assert (p["nsamples"], p["npulses"]) == (int(ref["nsamples"][0]), int(ref["npulses"][0]))
    assert p["delta_r"] == ref["delta_r"][0] and p["B_IF"] == ref["B_IF"][0]
    for k in ("k_r", "freq", "R_c", "pos", "phs", "k_y", "t"):
        assert np.array_equal(p[k], ref[k]), k
    sandia = os.path.join(REF_DATA, "Sandia/npy")
    for f in os.listdir(sandia):
        os.symlink(os.path.join(sandia, f), os.path.join(tmp_path, f))
    ref = {f[:-4]: np.load(os.path.join(sandia, f)) for f in os.listdir(sandia)}
    ns, np_ = int(ref["nsamples"][0]), int(ref["npulses"][0])
    np.save(os.path.join(tmp_path, "phs.npy"), np.ones((np_, ns), np.complex64))
    p = pl.load(str(tmp_path))
    assert (p["nsamples"], p["npulses"]) == (ns, np_)
    assert np.array_equal(p["k_r"], ref["k_r"].astype(np.float32))   # stored f64: downcast
    for k in ("R_c", "pos", "n_hat"):
        assert np.array_equal(p[k], ref[k]), k
    assert np.array_equal(p["t"], ref["t"]) and p["chirprate"] == ref["chirprate"][0]


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 0, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])
    rows = np.frombuffer(raw, np.uint8).reshape(h, w + 1)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:]


def test_png_writer_round_trip(tmp_path):
    img = np.random.default_rng(5).integers(0, 256, (37, 53), dtype=np.uint8)
    path = str(tmp_path / "x.png")
    _platform().write_png_gray8(path, img)
    assert np.array_equal(_read_png(path), img)
    with pytest.raises(_platform().PlatformError):
        _platform().write_png_gray8(str(tmp_path / "no" / "dir.png"), img)


def test_struct_layout_matches_header(tmp_path):
    pl = _platform()
    src = tmp_path / "l.c"
    fields = [f[0] for f in pl.PlatformStruct._fields_]
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "sarbp_host.h"\nint main(void){'
                   + 'printf("%zu\\n", sizeof(sarbp_platform));'
                   + "".join(f'printf("%zu\\n", offsetof(sarbp_platform, {f}));' for f in fields)
                   + "return 0;}")
    exe = tmp_path / "l"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(pl.PlatformStruct)
    assert vals[1:] == [getattr(pl.PlatformStruct, f).offset for f in fields]


def _run_sarbp(*args):
    return subprocess.run([_platform().SARBP_PATH, *args], capture_output=True, text=True,
                          timeout=120)


def test_cli_options_and_constraints(tmp_path):
    """sarbp.cpp:53-181: help, required options, the value constraints and schedule names."""
    r = _run_sarbp("-h")
    assert r.returncode == 0 and "--platform-dir=DIR" in r.stdout and "--num-iters=INT" in r.stdout
    r = _run_sarbp()
    assert r.returncode == 1 and "Missing required parameters" in r.stderr
    base = ["-p", str(tmp_path), "-o", str(tmp_path / "o.png")]
    r = _run_sarbp(*base)  # default dB range 0..0
    assert r.returncode == 1 and "Constraint failed: dB_min < dB_max" in r.stderr
    for opt, msg in (("-t0", "taylor >= 1"), ("-u0", "upsample >= 1"), ("-r0", "res-factor > 0"),
                     ("-a-1", "aspect > 0")):
        r = _run_sarbp(*base, "-d", "-30", "-D", "0", opt)
        assert r.returncode == 1 and msg in r.stderr, opt
    r = _run_sarbp(*base, "-d", "-30", "-D", "0", "-s", "cpu_distributed")
    assert r.returncode == 1 and "distributed support" in r.stderr
    r = _run_sarbp(*base, "-d", "-30", "-D", "0", "-s", "nope")
    assert r.returncode == 1 and "Unknown schedule: nope" in r.stderr
    r = _run_sarbp(*base, "--db-min=-30", "--db-max=0", "--schedule=ritsar")
    assert r.returncode == 1 and "not part of this library" in r.stderr


def test_cli_reports_loader_errors(tmp_path):
    _minimal(tmp_path, pos=np.zeros((10, 2), np.float32))
    r = _run_sarbp("-p", str(tmp_path), "-o", str(tmp_path / "o.png"), "-d", "-30", "-D", "0")
    assert r.returncode == 1 and "Bad shape: pos" in r.stderr
