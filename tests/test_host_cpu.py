"""CPU checks of the C++ drop-in's host logic (no GPU needed): XML config
tags (ophgen/src/ophGen.cpp:244-327, ophPointCloud.cpp:109-138) and PLY
loading semantics (Openholo/src/PLYparser.cpp:142-449), through the
ophpc_cli driver built from openholo_amd/host."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "openholo_amd", "bin", "ophpc_cli")
GOLD = os.path.join(ROOT, "tests", "golden")
VERTEX = np.dtype([("pos", "<f8", (3,)), ("color", "<f8", (3,)), ("phase", "<f8")])


def run_cli(*args, check=True):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    if check and r.returncode != 0:
        raise AssertionError(f"ophpc_cli failed: {r.stdout}\n{r.stderr}")
    return r


def parse_ascii_ply_reference_rules(path):
    """x, y, z as decimal doubles; colour = float32(int)/255 (single precision);
    phase rounded to float32; 0.5 colours when there is no 'color' element."""
    lines = open(path).read().splitlines()
    end = lines.index("end_header")
    header = lines[:end]
    has_color = any(l.startswith("element color") for l in header)
    elems = []
    for l in header:
        t = l.split()
        if t[0] == "element":
            elems.append([t[1], int(t[2]), []])
        elif t[0] == "property":
            elems[-1][2].append(t[-1])
    rows = lines[end + 1:]
    out = []
    pos = 0
    for name, count, props in elems:
        for k in range(count):
            vals = rows[pos].split()
            pos += 1
            if name != "vertex":
                continue
            d = dict(zip(props, vals))
            col = [np.float32(np.float32(int(d[c])) / np.float32(255.0)) for c in ("red", "green", "blue")]
            out.append(((float(d["x"]), float(d["y"]), float(d["z"])),
                        tuple(float(c) for c in col) if has_color else (0.5, 0.5, 0.5),
                        float(np.float32(float(d["phase"]))) if "phase" in d else 0.0))
    v = np.zeros(len(out), VERTEX)
    for i, (p, c, ph) in enumerate(out):
        v[i] = (p, c, ph)
    return v


def test_config_tags(tmp_path):
    r = run_cli("-c", os.path.join(GOLD, "Generation_PointCloud_RGB.xml"), "-i",
                os.path.join(GOLD, "pointcloud_1470.ply"), "--no-generate", "--dump-config")
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert cfg["pn"] == [1920, 1080]
    assert cfg["pp"] == [10e-6, 10e-6]
    assert cfg["lambdas"] == [638e-9, 520e-9, 450e-9]
    assert cfg["scale"] == [0.01, 0.01, 0.01]
    assert cfg["distance"] == 0.5
    assert (cfg["rotate"], cfg["merge"], cfg["flip"]) == (1, 1, 0)
    assert cfg["field_length"] == 1.0
    assert cfg["points"] == 1470
    r = run_cli("-c", os.path.join(GOLD, "Dice_PointCloud.xml"), "-i", os.path.join(GOLD, "pointcloud_1470.ply"),
                "--no-generate", "--dump-config")
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert cfg["pn"] == [1920, 1080] and cfg["pp"] == [4.3e-6, 4.3e-6] and cfg["lambdas"] == [532e-9]
    assert cfg["distance"] == 0.1 and cfg["rotate"] == 0


def test_config_missing_tag_fails(tmp_path):
    src = open(os.path.join(GOLD, "Dice_PointCloud.xml")).read()
    bad = tmp_path / "bad.xml"
    bad.write_text(src.replace("<SLM_PixelNumX>1920</SLM_PixelNumX>", ""))
    r = run_cli("-c", str(bad), "-i", os.path.join(GOLD, "pointcloud_1470.ply"), "--no-generate", check=False)
    assert r.returncode == 1
    bad.write_text(src.replace("<Distance>0.1</Distance>", ""))
    assert run_cli("-c", str(bad), "-i", os.path.join(GOLD, "pointcloud_1470.ply"), "--no-generate",
                   check=False).returncode == 1
    txt = tmp_path / "cfg.txt"
    txt.write_text(src)
    assert run_cli("-c", str(txt), "-i", os.path.join(GOLD, "pointcloud_1470.ply"), "--no-generate",
                   check=False).returncode == 1


def test_ascii_ply_matches_reference_rules(tmp_path):
    out = tmp_path / "p.bin"
    run_cli("-c", os.path.join(GOLD, "Dice_PointCloud.xml"), "-i", os.path.join(GOLD, "pointcloud_1470.ply"),
            "--no-generate", "--dump-points", str(out))
    got = np.fromfile(out, VERTEX)
    want = parse_ascii_ply_reference_rules(os.path.join(GOLD, "pointcloud_1470.ply"))
    assert len(got) == 1470
    assert np.array_equal(got["pos"], want["pos"])
    assert np.array_equal(got["color"], want["color"])
    assert np.array_equal(got["phase"], want["phase"])


def _write_binary_ply(path, verts, big=False, with_color_elem=True, extra=True):
    e = ">" if big else "<"
    hdr = ["ply", f"format binary_{'big' if big else 'little'}_endian 1.0"]
    if with_color_elem:
        hdr += ["element color 1", "property int channel"]
    hdr += [f"element vertex {len(verts)}", "property float x", "property float y", "property float z"]
    if extra:
        hdr += ["property float nx"]
    hdr += ["property uchar red", "property uchar green", "property uchar blue", "property float phase",
            "element face 1", "property list uchar int vertex_indices", "end_header"]
    body = b""
    if with_color_elem:
        body += struct.pack(e + "i", 3)
    for (x, y, z, r, g, b, ph) in verts:
        body += struct.pack(e + "fff", x, y, z)
        if extra:
            body += struct.pack(e + "f", 9.0)
        body += struct.pack(e + "BBBf", r, g, b, ph)
    body += struct.pack(e + "B", 3) + struct.pack(e + "iii", 0, 1, 2)
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode() + body)


@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("color_elem", [True, False])
def test_binary_ply(tmp_path, big, color_elem):
    verts = [(0.25, -1.5, 3.0, 10, 200, 255, 1.25), (-2.0, 0.125, 7.5, 0, 1, 2, -0.5)]
    ply = tmp_path / "b.ply"
    _write_binary_ply(ply, verts, big=big, with_color_elem=color_elem)
    out = tmp_path / "p.bin"
    run_cli("-c", os.path.join(GOLD, "Dice_PointCloud.xml"), "-i", str(ply), "--no-generate",
            "--dump-points", str(out))
    got = np.fromfile(out, VERTEX)
    assert len(got) == 2
    for g, (x, y, z, r, gg, b, ph) in zip(got, verts):
        assert tuple(g["pos"]) == (x, y, z)
        want = [float(np.float32(c) / np.float32(255)) for c in (r, gg, b)] if color_elem else [0.5] * 3
        assert list(g["color"]) == want
        assert g["phase"] == ph


def test_missing_ply_fails(tmp_path):
    r = run_cli("-c", os.path.join(GOLD, "Dice_PointCloud.xml"), "-i", str(tmp_path / "none.ply"),
                "--no-generate", check=False)
    assert r.returncode == 1


@pytest.mark.parametrize("count", ["-5", "4000000000"])
def test_ply_impossible_vertex_count_rejected(tmp_path, count):
    """A header count the file cannot hold (or a negative one) fails the load
    instead of allocating for it."""
    p = tmp_path / "bad.ply"
    p.write_text("ply\nformat ascii 1.0\nelement vertex %s\nproperty float x\nproperty float y\n"
                 "property float z\nend_header\n0 0 0\n1 1 1\n" % count)
    r = run_cli("-c", os.path.join(GOLD, "Dice_PointCloud.xml"), "-i", str(p), "--no-generate", check=False)
    assert r.returncode == 1


def test_truncated_binary_ply_rejected(tmp_path):
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex 50\nproperty float x\nproperty float y\n"
            "property float z\nend_header\n").encode()
    p = tmp_path / "short.ply"
    p.write_bytes(head + struct.pack("<30f", *range(30)))     # 10 of the 50 records
    r = run_cli("-c", os.path.join(GOLD, "Dice_PointCloud.xml"), "-i", str(p), "--no-generate", check=False)
    assert r.returncode == 1


@pytest.mark.parametrize("tag,value", [("SLM_PixelNumX", "-5"), ("SLM_PixelNumX", "2000000"),
                                       ("SLM_WaveNum", "0"), ("SLM_WaveNum", "4"), ("SLM_PixelPitchX", "nan")])
def test_config_out_of_range_rejected(tmp_path, tag, value):
    """Pixel numbers, wavelength count and pitch are checked before any buffer
    is sized from them (a channel indexes Vertex::color[3])."""
    text = open(os.path.join(GOLD, "Dice_PointCloud.xml")).read()
    start = text.index("<%s>" % tag) + len(tag) + 2
    end = text.index("</%s>" % tag)
    cfg = tmp_path / "bad.xml"
    cfg.write_text(text[:start] + value + text[end:])
    r = run_cli("-c", str(cfg), "-i", os.path.join(GOLD, "pointcloud_1470.ply"), "--no-generate", check=False)
    assert r.returncode == 1


def _ohc_bytes(fields, pp, lambdas, arrange=2, dtype=10, nw_header=None, px_header=None, offset_delta=0,
               cut=0):
    """An OHC file (Openholo/src/ImgCodecDefine.h:137-206): 24 B file header,
    65 B field-info header, wavelength table, then the real planes and the
    imaginary planes; arrange 2 = one plane per wavelength, 1 = wavelengths
    interleaved per pixel; dtype 10 = FP64, 9 = FP32."""
    nw, ny, nx = fields.shape
    es = 8 if dtype == 10 else 4
    header_size = 65 + 8 * nw
    field_size = nx * ny * nw * 2 * es
    off = 24 + header_size + offset_delta
    fh = struct.pack("<2sQBBIII", b"OH", 24 + header_size + field_size, 3, 0, 0, 0, off)
    pnx, pny = px_header or (nx, ny)
    fi = struct.pack("<IIIddBIBBBBBBBddQB", header_size, pnx, pny, pp[0], pp[1], 1,
                     nw if nw_header is None else nw_header, 2, arrange, 1, dtype, 1, 2, 0, -1.0, 1.0,
                     field_size, 0)
    wl = struct.pack("<%dd" % nw, *lambdas)
    ft = "<f8" if dtype == 10 else "<f4"
    planes = []
    for part in (fields.real, fields.imag):
        if arrange == 2:
            planes += [part[w].astype(ft).tobytes() for w in range(nw)]
        else:
            planes.append(np.stack([part[w].ravel() for w in range(nw)], axis=1).astype(ft).tobytes())
    data = fh + fi + wl + b"".join(planes)
    return data[:len(data) - cut] if cut else data


@pytest.mark.parametrize("arrange,dtype", [(2, 10), (1, 9)])
def test_load_ohc(tmp_path, arrange, dtype):
    """Openholo::loadAsOhc: per-wavelength planes (FP64) and per-pixel
    interleaved wavelengths (FP32); pitch, wavelengths and pixel numbers from
    the headers."""
    rng = np.random.default_rng(arrange)
    fields = rng.normal(size=(2, 5, 7)) + 1j * rng.normal(size=(2, 5, 7))
    p = tmp_path / "f.ohc"
    p.write_bytes(_ohc_bytes(fields, (8e-6, 6e-6), (638e-9, 532e-9), arrange, dtype))
    r = run_cli("--load-ohc", str(p), "--dump-field", str(tmp_path / "f.bin"))
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert cfg["pn"] == [7, 5] and cfg["pp"] == [8e-6, 6e-6] and cfg["lambdas"] == [638e-9, 532e-9]
    got = np.fromfile(tmp_path / "f.bin", np.complex128).reshape(2, 5, 7)
    want = fields if dtype == 10 else (fields.real.astype(np.float32) + 1j * fields.imag.astype(np.float32))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bad", [dict(nw_header=1_000_000_000), dict(px_header=(1 << 21, 5)),
                                 dict(px_header=(70000, 70000)), dict(offset_delta=10_000), dict(cut=16)])
def test_load_ohc_rejects_inconsistent_headers(tmp_path, bad):
    """Header sizes are checked against the file before any buffer is sized."""
    fields = np.ones((1, 5, 7), np.complex128)
    p = tmp_path / "bad.ohc"
    p.write_bytes(_ohc_bytes(fields, (8e-6, 8e-6), (532e-9,), **bad))
    r = run_cli("--load-ohc", str(p), check=False)
    assert r.returncode == 1


def test_set_pixel_number_without_initialize_is_refused(tmp_path):
    """setPixelNumber alone enlarges the configured hologram but leaves the
    host planes (complex_H, encoded, normalised) at the size initialize()
    gave them.  generateHologram, encoding and normalize refuse to run
    rather than write past those planes; each reports it through
    lastError().  Runs without a device: the check comes first."""
    r = run_cli("-c", os.path.join(GOLD, "Dice_PointCloud.xml"), "-i", os.path.join(GOLD, "pointcloud_1470.ply"),
                "-p", "1", "-f", "0", "-o", str(tmp_path / "R"), "--set-pixel-number", "2500,1500", check=False)
    assert r.returncode == 3, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert any(l.startswith("generateHologram: 0 s") and "host buffers do not match" in l for l in lines)
    assert any(l.startswith("encoding lastError: encoding: host buffers do not match") for l in lines)
    assert any(l.startswith("normalize lastError: normalize: host buffers do not match") for l in lines)
