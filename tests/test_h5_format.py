"""The self-written HDF5 writer/reader of libqss.so (Mesh::Save / Mesh::Load's
file format, save.cpp:178-835, myhdfshell.cpp) against real HDF5 readers and
writers.

Pins (each skips when its tool is absent from the image):
  * our writer -> our reader: every element kind, bit-exact;
  * our writer -> h5dump (HDF5 1.10 command-line tool): structure, types and
    values as that independent reader decodes them;
  * our writer -> libhdf5 (tests/native/h5_interop.c, H5Dread into native
    types): bit-exact values of F64, I64 and SavedIon datasets;
  * libhdf5 writer (the call sequence of the reference's Save) -> our reader:
    values of closed forms, including a root group large enough that libhdf5
    splits it over several symbol-table nodes;
  * refusals: a non-HDF5 file, a truncated file, a missing dataset, bad
    declarations.
No GPU is used."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from tertia_amd import h5
from tertia_amd._lib import QssError

HERE = os.path.dirname(os.path.abspath(__file__))
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)
HDF5_PREFIX = "/opt/conda"


def sample(seed=3, nsets=21):
    rng = np.random.default_rng(seed)
    ions = np.zeros(7, dtype=h5.SAVED_ION)
    parts = np.zeros(4, dtype=h5.SAVED_PARTICLE)
    for a in (ions, parts):
        for f in ("x", "y", "z", "px", "py", "pz", "weight", "q2m"):
            a[f] = rng.standard_normal(a.size)
        a["cell"] = rng.integers(-5, 10 ** 6, a.size)
    ions["i_z"] = rng.integers(0, 3, ions.size)
    ions["sort"] = 1
    d = {"meshinfo": rng.integers(-2 ** 31, 2 ** 31 - 1, 16).astype(np.int32),
         "phase": np.array([0.1 + 0.2]),
         "X": rng.standard_normal(9),
         "ex": rng.standard_normal((3, 5, 4)),
         "jz": rng.standard_normal((2, 6)),
         "Nparticles_at_PE_0": np.array([0, 7, 2 ** 40], dtype=np.int64),
         "Ions_of_sort_01_at_PE_0": ions,
         "Electrons_at_PE_0": parts}
    for i in range(nsets - len(d)):
        d["Dens%d" % i] = rng.standard_normal((2, 3, 2)) * 10.0 ** (i - 5)
    return d


def same(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def test_round_trip_every_kind(tmp_path):
    d = sample()
    path = str(tmp_path / "t.h5")
    h5.write(path, d)
    assert h5.names(path) == sorted(d, key=lambda s: s.encode())
    for name, v in d.items():
        kind, shape = h5.info(path, name)
        assert shape == v.shape, name
        r = h5.read(path, name)
        assert r.dtype == v.dtype and same(r, v), name
    # conversions: int32 -> float, int64 -> int32 of small values
    assert np.array_equal(h5.read(path, "meshinfo", h5.F64), d["meshinfo"].astype(float))
    assert np.array_equal(h5.read(path, "Nparticles_at_PE_0", h5.I32)[:2], [0, 7])
    # a particle compound read as the smaller record, member by member
    p = h5.read(path, "Ions_of_sort_01_at_PE_0", h5.PARTICLE)
    for f in h5.SAVED_PARTICLE.names:
        assert np.array_equal(p[f], d["Ions_of_sort_01_at_PE_0"][f]), f


def test_refusals(tmp_path):
    d = sample(nsets=9)
    path = str(tmp_path / "t.h5")
    h5.write(path, d)
    with pytest.raises(QssError, match="no dataset"):
        h5.read(path, "nope")
    with pytest.raises(QssError, match="not an HDF5 file"):
        open(tmp_path / "x.h5", "wb").write(b"plain text, not hdf5" * 10)
        h5.names(str(tmp_path / "x.h5"))
    raw = open(path, "rb").read()
    cut = str(tmp_path / "cut.h5")
    open(cut, "wb").write(raw[: len(raw) - 64])
    with pytest.raises(QssError, match="truncated|outside"):
        h5.read(cut, "jz")
    with pytest.raises(QssError, match="compound"):
        h5.read(path, "ex", h5.ION)
    with pytest.raises(QssError, match="floating-point"):
        h5.read(path, "ex", h5.I64)
    with pytest.raises(ValueError):
        h5.write(str(tmp_path / "r4.h5"), {"a": np.zeros((1, 1, 1, 1))})
    with pytest.raises(QssError, match="'/'"):
        h5.write(str(tmp_path / "sl.h5"), {"a/b": np.zeros(2)})


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not in this image")
def test_h5dump_reads_our_file(tmp_path):
    d = sample()
    path = str(tmp_path / "t.h5")
    h5.write(path, d)
    r = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    hdr = r.stdout
    found = re.findall(r'DATASET "([^"]+)"', hdr)
    assert found == sorted(d, key=lambda s: s.encode())
    assert re.search(r'DATASET "ex" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE '
                     r'\{ \( 3, 5, 4 \) / \( 3, 5, 4 \) \}', hdr)
    assert re.search(r'DATASET "meshinfo" \{\s*DATATYPE\s+H5T_STD_I32LE', hdr)
    assert re.search(r'DATASET "Nparticles_at_PE_0" \{\s*DATATYPE\s+H5T_STD_I64LE', hdr)
    comp = re.search(r'DATASET "Ions_of_sort_01_at_PE_0" \{\s*DATATYPE\s+H5T_COMPOUND \{(.*?)\}',
                     hdr, re.S).group(1)
    assert re.findall(r'H5T_\w+ "(\w+)";', comp) == list(h5.SAVED_ION.names)
    # values as h5dump prints them (%.17g for doubles)
    r = subprocess.run([H5DUMP, "-d", "/ex", "-w", "0", "-m", "%.17g", "-y", path],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    body = r.stdout.split("DATA {", 1)[1]
    vals = np.array([float(x) for x in re.findall(r"-?\d+\.?\d*(?:e[-+]?\d+)?", body)])
    assert np.array_equal(vals, d["ex"].ravel())


@pytest.fixture(scope="module")
def interop(tmp_path_factory):
    inc = os.path.join(HDF5_PREFIX, "include", "hdf5.h")
    if not os.path.exists(inc) or shutil.which("gcc") is None:
        pytest.skip("no libhdf5 development files in this image")
    exe = str(tmp_path_factory.mktemp("h5i") / "h5_interop")
    lib = os.path.join(HDF5_PREFIX, "lib")
    r = subprocess.run(["gcc", "-O1", os.path.join(HERE, "native", "h5_interop.c"),
                        "-I", os.path.join(HDF5_PREFIX, "include"), "-L", lib, "-lhdf5",
                        "-Wl,-rpath," + lib, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("cannot link against libhdf5: " + r.stderr[-300:])
    return exe


def run_interop(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    dims = tuple(int(x) for x in lines[0].split()[1:])
    return dims, lines[1:]


def test_libhdf5_reads_our_file(tmp_path, interop):
    d = sample()
    path = str(tmp_path / "t.h5")
    h5.write(path, d)
    for name in ("ex", "jz", "X", "phase") + tuple(k for k in d if k.startswith("Dens")):
        dims, lines = run_interop(interop, "f64", path, name)
        assert dims == d[name].shape
        assert np.array_equal(np.array([float(x) for x in lines]), d[name].ravel()), name
    dims, lines = run_interop(interop, "i64", path, "Nparticles_at_PE_0")
    assert [int(x) for x in lines] == d["Nparticles_at_PE_0"].tolist()
    dims, lines = run_interop(interop, "f64", path, "meshinfo")   # I32 read as double
    assert np.array_equal(np.array([float(x) for x in lines]), d["meshinfo"].astype(float))
    dims, lines = run_interop(interop, "ion", path, "Ions_of_sort_01_at_PE_0")
    ions = d["Ions_of_sort_01_at_PE_0"]
    assert dims == ions.shape
    for rec, line in zip(ions, lines):
        f = line.split()
        assert int(f[0]) == rec["cell"] and int(f[9]) == rec["i_z"] and int(f[10]) == rec["sort"]
        got = np.array([float(x) for x in f[1:9]])
        want = np.array([rec[k] for k in ("x", "y", "z", "px", "py", "pz", "weight", "q2m")])
        assert np.array_equal(got, want)


def test_our_reader_reads_libhdf5_file(tmp_path, interop):
    path = str(tmp_path / "ref.h5")
    run_interop_write = subprocess.run([interop, "write", path], capture_output=True, text=True)
    assert run_interop_write.returncode == 0, run_interop_write.stderr
    fields = ["ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "dens", "Dens0", "Dens1"]
    assert set(h5.names(path)) == set(fields) | {"meshinfo", "phase", "Ions_of_sort_01_at_PE_0",
                                                 "Nparticles_at_PE_0"}
    assert np.array_equal(h5.read(path, "meshinfo"), 3 * np.arange(16) - 7)
    assert h5.read(path, "phase").tolist() == [2.75]
    k, j, i = np.meshgrid(np.arange(5), np.arange(4), np.arange(6), indexing="ij")
    for fi, name in enumerate(fields):
        assert np.array_equal(h5.read(path, name), i + 10.0 * j + 100.0 * k + 0.25 * fi), name
    p = h5.read(path, "Ions_of_sort_01_at_PE_0")
    r = np.arange(37)
    assert np.array_equal(p["cell"], 1000 + r) and np.array_equal(p["x"], r + 0.5)
    assert np.array_equal(p["y"], -r.astype(float)) and np.array_equal(p["z"], 0.125 * r)
    assert np.array_equal(p["px"], 1e-3 * r) and np.array_equal(p["weight"], 1.0 / (r + 1))
    assert (p["py"] == 2).all() and (p["pz"] == -3).all() and (p["q2m"] == -1).all()
    assert (p["i_z"] == 1).all() and (p["sort"] == 1).all()
    assert h5.read(path, "Nparticles_at_PE_0").tolist() == [0, 37]
