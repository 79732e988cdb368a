"""Structured volume readers of the headless host layer (VolumeReader: .raw, .syn,
.nrrd/.nhrd, .dat) and the dataset lists (DataProvider: the extension-less
"<path> <name>" list and the CSV list), checked through cvr_cli --check-inputs against
volumes built in numpy.  The two dataset lists are the reference's own files
(data/volume_list.csv, data/#list_structured_datasets); the Bonsai voxels they name
are not part of the reference, so a seeded volume of the same size stands in."""
import shutil
from pathlib import Path

import numpy as np
import pytest

from test_host_layer import parse_check, run

GOLD = Path(__file__).parent / "golden" / "datasets"


def wsum(vol):
    """sum over voxels i (x fastest) of (i + 1) * value, modulo 2^64 (what the CLI prints)."""
    v = np.ascontiguousarray(vol).reshape(-1).astype(np.uint64)
    return int((np.arange(1, v.size + 1, dtype=np.uint64) * v).sum(dtype=np.uint64))


def check(*args):
    out = run(*args, "--check-inputs").stdout
    kv = parse_check(out)
    line = next(l for l in out.splitlines() if l.startswith("volume "))
    dims = line.split()[2]
    return out, kv, dims


def bonsai_like(rng_seed=7):
    rng = np.random.default_rng(rng_seed)
    return rng.integers(0, 256, (256, 256, 256), dtype=np.uint8)  # (z, y, x)


def test_csv_dataset_list_fixture(tmp_path):
    shutil.copy(GOLD / "volume_list.csv", tmp_path / "volume_list.csv")
    (tmp_path / "raw").mkdir()
    vol = bonsai_like()
    vol.tofile(tmp_path / "raw" / "Bonsai.1.256x256x256.raw")
    out, kv, dims = check("--datasets", tmp_path / "volume_list.csv", "--dataset", "Bonsai")
    assert "datasets [Bonsai]" in out
    assert dims == "256x256x256" and kv["bytes"] == "1" and kv["scale"] == "1,1,1"
    assert int(kv["wsum"]) == wsum(vol)
    # by index, too
    assert parse_check(run("--datasets", tmp_path / "volume_list.csv", "--dataset", "0", "--check-inputs").stdout)[
        "wsum"] == kv["wsum"]


def test_plain_dataset_list_fixture(tmp_path):
    shutil.copy(GOLD / "#list_structured_datasets", tmp_path / "#list_structured_datasets")
    (tmp_path / "raw").mkdir()
    vol = bonsai_like(8)
    vol.tofile(tmp_path / "raw" / "Bonsai.1.256x256x256.raw")
    out, kv, dims = check("--datasets", tmp_path / "#list_structured_datasets", "--dataset", "Structured Bonsai")
    assert "datasets [Structured Bonsai]" in out
    assert dims == "256x256x256" and int(kv["wsum"]) == wsum(vol)


def test_csv_list_spacing_ushort_and_other_formats(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.integers(0, 65536, (5, 6, 7), dtype=np.uint16)
    (tmp_path / "vols").mkdir()
    a.astype("<u2").tofile(tmp_path / "vols" / "a.raw")
    b = rng.integers(0, 256, (3, 4, 5), dtype=np.uint8)
    b.tofile(tmp_path / "vols" / "b.bin")
    (tmp_path / "vols" / "b.dat").write_text("ObjectFileName: b.bin\nResolution: 5 4 3\nSliceThickness: 1 1 2.5\n"
                                             "Format: UCHAR\n")
    (tmp_path / "list.csv").write_text("Name,Source,File,Size,Spacing,Type\n"
                                       "A,vols,a.raw,7x6x5,0.5x1x2,ushort\n"
                                       "B,vols,b.dat,0x0x0,1x1x1,uchar\n")
    # no "sep=" line: the default separator is ';', so a comma list has one column
    r = run("--datasets", tmp_path / "list.csv", "--dataset", "A", "--check-inputs", check=False)
    assert r.returncode == 1 and "column" in r.stderr
    (tmp_path / "list.csv").write_text("sep=,\n" + (tmp_path / "list.csv").read_text())
    out, kv, dims = check("--datasets", tmp_path / "list.csv", "--dataset", "A")
    assert "datasets [A] [B]" in out
    assert dims == "7x6x5" and kv["bytes"] == "2" and kv["scale"] == "0.5,1,2"
    assert int(kv["wsum"]) == wsum(a)
    out, kv, dims = check("--datasets", tmp_path / "list.csv", "--dataset", "B")
    assert dims == "5x4x3" and kv["scale"] == "1,1,2.5" and int(kv["wsum"]) == wsum(b)
    r = run("--datasets", tmp_path / "list.csv", "--dataset", "C", "--check-inputs", check=False)
    assert r.returncode == 1 and "no dataset" in r.stderr


def syn_volume(w, h, d, records):
    v = np.zeros((d, h, w), dtype=np.uint8)
    for r in records:
        if r[0] == 1:
            _, x0, y0, z0, x1, y1, z1, val = r
            v[z0:z1, y0:y1, x0:x1] = val % 256
        else:
            _, x, y, z, val = r
            v[z, y, x] = val % 256
    return v


def test_syn_reader(tmp_path):
    recs = [(1, 0, 0, 0, 4, 3, 2, 200), (0, 5, 1, 3, 77), (1, 2, 1, 1, 6, 4, 4, 300), (1, 3, 3, 3, 3, 4, 4, 9),
            (2, 0, 3, 3, 5)]
    p = tmp_path / "boxes.syn"
    p.write_text("6 4 5\n" + "\n".join(" ".join(map(str, r)) for r in recs) + "\n")
    out, kv, dims = check("--volume", p)
    ref = syn_volume(6, 4, 5, recs)
    assert dims == "6x4x5" and kv["bytes"] == "1"
    assert int(kv["wsum"]) == wsum(ref)
    assert float(kv["sample(0,0,0)"]) == pytest.approx(200 / 255, abs=1e-9)
    # a record outside the volume and a truncated record are errors, not memory writes
    (tmp_path / "bad.syn").write_text("2 2 2\n0 2 0 0 5\n")
    r = run("--volume", tmp_path / "bad.syn", "--check-inputs", check=False)
    assert r.returncode == 1 and "outside" in r.stderr
    (tmp_path / "trunc.syn").write_text("2 2 2\n1 0 0 0 2 2\n")
    r = run("--volume", tmp_path / "trunc.syn", "--check-inputs", check=False)
    assert r.returncode == 1 and "truncated" in r.stderr


@pytest.mark.parametrize("ext", ["nrrd", "nhrd"])
def test_nrrd_detached_and_attached(tmp_path, ext):
    rng = np.random.default_rng(4)
    v = rng.integers(0, 65536, (4, 3, 6), dtype=np.uint16)
    v.astype("<u2").tofile(tmp_path / "data.raw")
    hdr = ("NRRD0004\n# a comment\ntype: unsigned short\ndimension: 3\nspace: left-posterior-superior\n"
           "sizes: 6 3 4\nspace directions: (1,0,0) (0,1,0) (0,0,1)\nkinds: domain domain domain\n"
           "endian: little\nencoding: raw\ndata file: data.raw\n")
    (tmp_path / f"v.{ext}").write_text(hdr)
    out, kv, dims = check("--volume", tmp_path / f"v.{ext}")
    assert dims == "6x3x4" and kv["bytes"] == "2" and int(kv["wsum"]) == wsum(v)
    # attached: the data follows the header's empty line
    att = tmp_path / "att.nrrd"
    att.write_bytes(hdr.replace("data file: data.raw\n", "").encode() + b"\n" + v.astype("<u2").tobytes())
    out, kv, dims = check("--volume", att)
    assert int(kv["wsum"]) == wsum(v)


def test_nrrd_rejects_what_it_cannot_read(tmp_path):
    base = "NRRD0004\ntype: uint8\ndimension: 3\nsizes: 2 2 2\nencoding: raw\ndata file: d.raw\n"
    (tmp_path / "d.raw").write_bytes(bytes(8))
    for bad, msg in ((base.replace("encoding: raw", "encoding: gzip"), "raw encoding"),
                     (base.replace("type: uint8", "type: float"), "unsigned 8 or 16"),
                     (base.replace("dimension: 3", "dimension: 2").replace("sizes: 2 2 2", "sizes: 2 2"), "dimension 3"),
                     (base.replace("sizes: 2 2 2", "sizes: 2 2 3"), "short read"),
                     (base.replace("NRRD0004", "NOTNRRD"), "NRRD header")):
        (tmp_path / "x.nrrd").write_text(bad)
        r = run("--volume", tmp_path / "x.nrrd", "--check-inputs", check=False)
        assert r.returncode == 1 and msg in r.stderr, (bad, r.stderr)


def test_dat_reader(tmp_path):
    rng = np.random.default_rng(5)
    v = rng.integers(0, 65536, (3, 5, 4), dtype=np.uint16)
    (tmp_path / "sub").mkdir()
    v.astype("<u2").tofile(tmp_path / "sub" / "obj.raw")
    (tmp_path / "sub" / "m.dat").write_text("ObjectFileName:  obj.raw\nTaggedFileName: ---\nResolution:  4 5 3\n"
                                            "SliceThickness: 0.5 0.5 1.5\nFormat: USHORT\nNbrTags: 0\n"
                                            "ObjectType: TEXTURE_VOLUME_OBJECT\n")
    out, kv, dims = check("--volume", tmp_path / "sub" / "m.dat")
    assert dims == "4x5x3" and kv["bytes"] == "2" and kv["scale"] == "0.5,0.5,1.5"
    assert int(kv["wsum"]) == wsum(v)
    (tmp_path / "sub" / "f.dat").write_text("ObjectFileName: obj.raw\nResolution: 4 5 3\nFormat: FLOAT\n")
    r = run("--volume", tmp_path / "sub" / "f.dat", "--check-inputs", check=False)
    assert r.returncode == 1 and "UCHAR or USHORT" in r.stderr


def test_pvm_is_rejected(tmp_path):
    (tmp_path / "x.pvm").write_bytes(b"PVM\n")
    r = run("--volume", tmp_path / "x.pvm", "--check-inputs", check=False)
    assert r.returncode == 1 and "not supported" in r.stderr
