"""The FacesNMF image loader of the drop-in driver (gpumlib_amd/examples/faces_nmf.cpp,
load_pgm) against the oracle's restatement of LoadImage (FacesNMFDlg.cpp:409-472), bit for
bit, on synthetic P5 files and on the CBCL images the demo trains on.  CPU only: the
driver's --load-only mode loads and writes V and touches no GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpumlib_amd", "bin", "faces_nmf")
FIXTURE = os.path.join(ROOT, "tests", "golden", "cbcl_faces.npz")
PIXELS = 361


def write_files(d, blobs):
    d.mkdir(exist_ok=True)
    for k, b in enumerate(blobs):
        (d / f"face{k + 1:05d}.pgm").write_bytes(b)


def load(d, count, equalize, standard, out):
    args = [EXE, "--pgm", str(d), "--images", str(count), "--load-only", "--dump-v", str(out)]
    args += ["--equalize"] if equalize else []
    args += ["--pgm-standard"] if standard else []
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return np.fromfile(out, np.float32).reshape(count, PIXELS)  # column k of V is row k here


def header(crlf=False):
    return b"P5\r\n19 19\r\n255\r\n" if crlf else b"P5\n19 19\n255\n"


def test_oracle_loader_known_answers(orc):
    """Pins the restatement by cases worked out by hand from FacesNMFDlg.cpp:409-472."""
    pix = bytes((7 * k + 3) % 256 for k in range(PIXELS))
    v = orc.pgm_column(header() + pix, equalize=False)
    # the reference reads the raster from the byte after "255": the newline (10) is pixel 0
    assert v[0] == np.float32(10) / np.float32(255) and v[1] == np.float32(pix[0]) / np.float32(255)
    assert np.array_equal(v[1:], np.frombuffer(pix[:-1], np.uint8).astype(np.float32) / np.float32(255))
    s = orc.pgm_column(header() + pix, equalize=False, standard=True)
    assert np.array_equal(s, np.frombuffer(pix, np.uint8).astype(np.float32) / np.float32(255))
    # CR LF header: "\r\n" are pixels 0 and 1
    c = orc.pgm_column(header(True) + pix, equalize=False)
    assert c[0] == np.float32(13) / np.float32(255) and c[1] == np.float32(10) / np.float32(255)
    # equalisation of a two-level image: the lower level maps to 0, the upper to 1
    two = bytes([40] * 100 + [200] * 261)
    e = orc.pgm_column(header() + two, equalize=True, standard=True)
    assert set(np.unique(e).tolist()) == {0.0, 1.0} and e[0] == 0.0 and e[-1] == 1.0
    # three levels a < b < c with counts 1, k, rest: b maps to k / (361 - 1)
    three = bytes([5] + [50] * 120 + [90] * 240)
    e = orc.pgm_column(header() + three, equalize=True, standard=True)
    assert e[0] == 0.0 and e[1] == np.float32(120) / np.float32(360) and e[-1] == 1.0
    # a constant image: pixels - cdfmin = 0, every level 0 / 0
    assert np.all(np.isnan(orc.pgm_column(header() + bytes([77] * PIXELS), equalize=True, standard=True)))
    # what LoadImage rejects
    assert orc.pgm_column(b"P2\n19 19\n255\n" + pix, False) is None
    assert orc.pgm_column(b"P5\n19 20\n255\n" + pix, False) is None
    assert orc.pgm_column(b"P5\n19 19\n65535\n" + pix, False) is None


@pytest.mark.parametrize("equalize", [False, True])
@pytest.mark.parametrize("standard", [False, True])
def test_loader_synthetic_files(native, orc, tmp_path, equalize, standard):
    rng = np.random.default_rng(5)
    blobs = [
        header() + rng.integers(0, 256, PIXELS, dtype=np.uint8).tobytes(),
        header(True) + rng.integers(0, 256, PIXELS, dtype=np.uint8).tobytes(),
        header() + bytes([10, 13] * 180 + [255]),             # raster bytes that are also line ends
        header() + bytes([0] * 180 + [255] * 181),            # the extremes
        header() + bytes([123] * PIXELS),                     # constant: equalisation gives NaN
        header(True) + bytes(range(256)) + bytes(range(105)),
    ]
    write_files(tmp_path / "img", blobs)
    V = load(tmp_path / "img", len(blobs), equalize, standard, tmp_path / "v.bin")
    for k, b in enumerate(blobs):
        ref = orc.pgm_column(b, equalize, standard)
        assert np.array_equal(V[k], ref, equal_nan=True), k


def test_loader_rejects_a_bad_file(native, tmp_path):
    write_files(tmp_path / "img", [header() + bytes(PIXELS), b"P5\n18 19\n255\n" + bytes(PIXELS)])
    p = subprocess.run([EXE, "--pgm", str(tmp_path / "img"), "--images", "2", "--load-only"], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 1 and "face00002.pgm" in p.stderr


def cbcl_blobs():
    z = np.load(FIXTURE)
    data, off = z["data"].tobytes(), z["offsets"]
    return [data[off[k]:off[k + 1]] for k in range(off.size - 1)]


@pytest.mark.parametrize("equalize", [False, True])
def test_loader_cbcl_images(native, orc, tmp_path, equalize):
    """All 2429 images the demo loads, as the reference tree holds them (LF and CR LF files)."""
    blobs = cbcl_blobs()
    assert len(blobs) == 2429
    write_files(tmp_path / "img", blobs)
    V = load(tmp_path / "img", len(blobs), equalize, False, tmp_path / "v.bin")
    ref = np.stack([orc.pgm_column(b, equalize) for b in blobs])
    assert np.array_equal(V, ref, equal_nan=True)
    assert not np.any(np.isnan(V))
