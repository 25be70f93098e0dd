"""On-disk formats (include/dft_io.h): field files, patch shards and the
result document.  Host code only, so these run without a GPU.

The reference has no file format on this path (it prints its results), so
the pins are the formats' own definitions: CRC-32 against zlib and the
standard check value, the binary header parsed here with `struct` from the
layout the header file documents, exact round trips, and corruption
detection."""
import ctypes as C
import json
import math
import os
import struct
import zlib

import numpy as np
import pytest

import dft_amd

FIELD_HDR = 256
PATCH_HDR = 128


@pytest.fixture(scope="module")
def grid():
    return dft_amd.Grid.make(9, 7, 5, -2.0, 2.0, -1.5, 1.5, -1.0, 1.0)


def test_crc32_matches_zlib_and_check_value():
    assert dft_amd.crc32(b"123456789") == 0xCBF43926  # the CRC-32/ISO-HDLC catalogue check value
    assert dft_amd.crc32(b"") == 0
    rng = np.random.default_rng(0)
    for n in list(range(0, 40)) + [1000, 4099, 65537]:
        buf = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert dft_amd.crc32(buf) == zlib.crc32(buf), n
        # every split continues to the same value (8-byte steps and byte tails mixed)
        for cut in {0, n // 3, n // 2, max(n - 1, 0)}:
            assert dft_amd.crc32(buf[cut:], dft_amd.crc32(buf[:cut])) == zlib.crc32(buf), (n, cut)


def _parse_field_header(raw):
    """The field header exactly as include/dft_io.h lays it out."""
    assert raw[:8] == b"DFTFLD01"
    hb, dtype = struct.unpack_from("<II", raw, 8)
    nx, ny, nz, ncols = struct.unpack_from("<4q", raw, 16)
    geo = struct.unpack_from("<9d", raw, 48)
    nbytes, = struct.unpack_from("<Q", raw, 120)
    crc, = struct.unpack_from("<I", raw, 128)
    name = raw[132:196].split(b"\0", 1)[0].decode()
    assert raw[196:252] == b"\0" * 56
    assert struct.unpack_from("<I", raw, 252)[0] == zlib.crc32(raw[:252])  # the header's own checksum
    return dict(hb=hb, dtype=dtype, n=(nx, ny, nz), ncols=ncols, geo=geo, nbytes=nbytes, crc=crc, name=name)


@pytest.mark.parametrize("dtype,ncols", [(np.float64, 1), (np.float32, 1), (np.float64, 3), (np.float32, 5)])
def test_field_round_trip_and_layout(tmp_path, grid, dtype, ncols):
    rng = np.random.default_rng(ncols)
    a = rng.standard_normal((ncols, grid.size)).astype(dtype)
    a[0, :4] = [0.0, -0.0, np.inf, np.nan]  # stored as bits, whatever they are
    data = a[0] if ncols == 1 else a
    path = tmp_path / "f.dftf"
    dft_amd.save_field(path, grid, data, name="psi_alpha")
    assert not os.path.exists(str(path) + ".tmp")
    raw = path.read_bytes()
    assert len(raw) == FIELD_HDR + a.nbytes
    h = _parse_field_header(raw)
    assert h["hb"] == FIELD_HDR and h["dtype"] == (dft_amd.F64 if dtype == np.float64 else dft_amd.F32)
    assert h["n"] == (9, 7, 5) and h["ncols"] == ncols and h["nbytes"] == a.nbytes
    assert h["geo"] == (grid.xmin, grid.xmax, grid.ymin, grid.ymax, grid.zmin, grid.zmax, grid.dx, grid.dy, grid.dz)
    assert h["name"] == "psi_alpha"
    # the payload is the device layout: column after column, x fastest
    assert raw[FIELD_HDR:] == a.tobytes() and h["crc"] == zlib.crc32(a.tobytes())
    g2, b, name = dft_amd.load_field(path)
    assert name == "psi_alpha" and b.dtype == dtype and b.shape == data.shape
    assert b.tobytes() == data.tobytes()
    assert bytes(g2) == bytes(grid)
    info = dft_amd.field_info(path)
    assert info["ncols"] == ncols and info["crc"] == h["crc"]


def test_field_rejects_corruption(tmp_path, grid):
    a = np.arange(grid.size, dtype=np.float64)
    path = tmp_path / "rho.dftf"
    dft_amd.save_field(path, grid, a, name="rho")
    raw = bytearray(path.read_bytes())
    # one flipped data bit: checksum mismatch (numerical-failure status)
    bad = bytearray(raw)
    bad[FIELD_HDR + 17] ^= 0x10
    (tmp_path / "bit.dftf").write_bytes(bytes(bad))
    with pytest.raises(dft_amd.DftError, match="checksum") as e:
        dft_amd.load_field(tmp_path / "bit.dftf")
    assert e.value.code == -2
    # truncated payload
    (tmp_path / "short.dftf").write_bytes(bytes(raw[:-8]))
    with pytest.raises(dft_amd.DftError) as e:
        dft_amd.load_field(tmp_path / "short.dftf")
    assert e.value.code == -2
    # not a field file, or a header whose size disagrees with its shape
    (tmp_path / "magic.dftf").write_bytes(b"XXXXXXXX" + bytes(raw[8:]))
    with pytest.raises(dft_amd.DftError, match="not a field file") as e:
        dft_amd.field_info(tmp_path / "magic.dftf")
    assert e.value.code == -1
    # a changed header byte fails the header's checksum, wherever it is (here: the grid's xmax)
    bad = bytearray(raw)
    bad[57] ^= 0x01
    (tmp_path / "hdr.dftf").write_bytes(bytes(bad))
    with pytest.raises(dft_amd.DftError, match="header checksum") as e:
        dft_amd.field_info(tmp_path / "hdr.dftf")
    assert e.value.code == -2
    bad = bytearray(raw)
    struct.pack_into("<q", bad, 40, 2)  # ncols 2 with one column's bytes, header checksum redone
    struct.pack_into("<I", bad, 252, zlib.crc32(bytes(bad[:252])))
    (tmp_path / "shape.dftf").write_bytes(bytes(bad))
    with pytest.raises(dft_amd.DftError, match="disagrees") as e:
        dft_amd.field_info(tmp_path / "shape.dftf")
    assert e.value.code == -1
    with pytest.raises(dft_amd.DftError) as e:
        dft_amd.field_info(tmp_path / "missing.dftf")
    assert e.value.code == -1


def test_field_argument_checks(tmp_path, grid):
    L = dft_amd.core_lib()
    a = np.zeros(grid.size)
    with pytest.raises(dft_amd.DftError, match="63"):
        dft_amd.save_field(tmp_path / "n.dftf", grid, a, name="x" * 64)
    with pytest.raises(ValueError):
        dft_amd.save_field(tmp_path / "n.dftf", grid, np.zeros(grid.size + 1))
    with pytest.raises(ValueError):
        dft_amd.save_field(tmp_path / "n.dftf", grid, np.zeros(grid.size, dtype=np.int32))
    dft_amd.save_field(tmp_path / "ok.dftf", grid, a)
    small = np.zeros(grid.size - 1)
    hdr = dft_amd.FieldHeader()
    rc = L.dft_field_read(os.fsencode(tmp_path / "ok.dftf"), C.byref(hdr), small.ctypes.data_as(C.c_void_p),
                          C.c_uint64(small.nbytes))
    assert rc == -1
    assert L.dft_field_write(None, C.byref(grid), 0, C.c_int64(1), b"", a.ctypes.data_as(C.c_void_p)) == -1
    assert L.dft_field_write(os.fsencode(tmp_path / "d.dftf"), C.byref(grid), 7, C.c_int64(1), b"",
                             a.ctypes.data_as(C.c_void_p)) == -1
    # a directory that does not exist: the write fails and leaves nothing behind
    with pytest.raises(dft_amd.DftError, match="cannot open"):
        dft_amd.save_field(tmp_path / "no" / "such.dftf", grid, a)


def _patches(count, P, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((count, 2, P ** 3)).astype(np.float32),
            rng.standard_normal((count, P ** 3)).astype(np.float32))


def test_patch_shard_round_trip_and_layout(tmp_path):
    P, count = 6, 11
    xin, tgt = _patches(count, P)
    path = tmp_path / "s.dftp"
    dft_amd.save_patch_shard(path, xin, tgt, target_scale=1, h=0.2)
    raw = path.read_bytes()
    assert len(raw) == PATCH_HDR + xin.nbytes + tgt.nbytes
    assert raw[:8] == b"DFTPAT01"
    hb, ch = struct.unpack_from("<II", raw, 8)
    p, n = struct.unpack_from("<2q", raw, 16)
    ib, tb = struct.unpack_from("<2Q", raw, 32)
    ci, ct = struct.unpack_from("<2I", raw, 48)
    ts, = struct.unpack_from("<i", raw, 56)
    h, = struct.unpack_from("<d", raw, 64)
    assert (hb, ch, p, n, ib, tb, ts, h) == (PATCH_HDR, 2, P, count, xin.nbytes, tgt.nbytes, 1, 0.2)
    assert raw[72:124] == b"\0" * 52 and struct.unpack_from("<I", raw, 124)[0] == zlib.crc32(raw[:124])
    assert raw[PATCH_HDR:PATCH_HDR + ib] == xin.tobytes() and raw[PATCH_HDR + ib:] == tgt.tobytes()
    assert (ci, ct) == (zlib.crc32(xin.tobytes()), zlib.crc32(tgt.tobytes()))
    a, b, hdr = dft_amd.load_patch_shard(path)
    assert a.tobytes() == xin.tobytes() and b.tobytes() == tgt.tobytes()
    assert hdr["patch"] == P and hdr["count"] == count and hdr["target_scale"] == 1 and hdr["h"] == 0.2
    a, b, _ = dft_amd.load_patch_shard(path, first=3, count=5)
    assert np.array_equal(a, xin[3:8]) and np.array_equal(b, tgt[3:8])
    with pytest.raises(dft_amd.DftError, match="beyond"):
        dft_amd.load_patch_shard(path, first=8, count=4)


def test_patch_shard_empty_and_corrupt(tmp_path):
    P = 4
    xin, tgt = _patches(0, P)
    dft_amd.save_patch_shard(tmp_path / "e.dftp", xin.reshape(0, 2, P ** 3), tgt.reshape(0, P ** 3))
    a, b, hdr = dft_amd.load_patch_shard(tmp_path / "e.dftp")
    assert hdr["count"] == 0 and a.shape == (0, 2, P ** 3) and b.shape == (0, P ** 3)
    xin, tgt = _patches(3, P, seed=2)
    dft_amd.save_patch_shard(tmp_path / "c.dftp", xin, tgt)
    raw = bytearray((tmp_path / "c.dftp").read_bytes())
    raw[-3] ^= 1  # a target byte
    (tmp_path / "c.dftp").write_bytes(bytes(raw))
    with pytest.raises(dft_amd.DftError, match="target checksum") as e:
        dft_amd.load_patch_shard(tmp_path / "c.dftp")
    assert e.value.code == -2
    # a partial read checksums its whole section too: the damaged target section fails even for a
    # range that does not hold the damaged byte
    with pytest.raises(dft_amd.DftError, match="target checksum") as e:
        dft_amd.load_patch_shard(tmp_path / "c.dftp", 0, 2)
    assert e.value.code == -2


def _result(smearing=0.0):
    cfg = dft_amd.ScfConfig()
    cfg.functional, cfg.charge, cfg.multiplicity, cfg.precision = dft_amd.PBE96, 1, 2, dft_amd.SCF_MIXED
    cfg.nstates, cfg.poisson_method, cfg.stop_rule, cfg.energy_tol = 12, dft_amd.POISSON_DIRECT, dft_amd.STOP_DE, 1e-5
    cfg.init_guess, cfg.davidson_tol, cfg.precond, cfg.mixer = dft_amd.INIT_HYDROGENIC, 1e-7, dft_amd.PRECOND_KINETIC_EXACT, 1
    cfg.smearing = smearing
    rep = dft_amd.ScfReport()
    rep.iterations, rep.converged, rep.first_de_iteration, rep.refine_iterations = 17, 1, 14, 3
    energies = [1.0 / 3.0, -7.123456789012345, 2.5e-17, -0.1, 1e300, -4.440892098500626e-16]
    for i, e in enumerate(energies):
        rep.energy[i] = e
    rep.energy_difference, rep.density_difference, rep.ks_residual = -3.1e-6, float("nan"), float("inf")
    rep.seconds_total, rep.seconds_to_first_de = 0.287, 0.25
    rep.happly_columns, rep.poisson_sweeps, rep.davidson_iterations = 2 ** 40 + 3, 0, 69
    rep.fermi_level[0], rep.fermi_level[1], rep.smearing_energy = -0.25, -0.2, -1e-4
    # 4 alpha + 3 beta electrons: nuclei Z 6 + 2 - charge 1
    nuc = np.array([[6.0, 0.0, 0.0, 0.1], [2.0, 1.5, -0.25, 0.0]])
    eig = np.array([-10.1, -0.5, -0.3, -0.3 + 1e-15, -9.9, -0.45, -0.2])
    return nuc, dict(config=cfg, report=rep, eigenvalues=eig, n_alpha=4, n_beta=3), energies


def test_result_json_round_trips_every_number(grid):
    nuc, res, energies = _result()
    doc = json.loads(dft_amd.scf_result_json(grid, nuc, res))
    assert doc["format"] == "dft-scf-result" and doc["version"] == 1
    assert doc["electrons"] == 7 and doc["charge"] == 1
    assert doc["grid"]["n"] == [9, 7, 5] and doc["grid"]["spacing"] == [grid.dx, grid.dy, grid.dz]
    assert doc["nuclei"][1] == {"Z": 2.0, "position": [1.5, -0.25, 0.0]}
    c = doc["config"]
    assert (c["functional"], c["precision"], c["poisson"], c["stopRule"], c["preconditioner"], c["mixer"]) == \
        ("PBE96", "mixed", "direct", "energy", "kinetic_exact", "pulay")
    scf = doc["scf"]
    names = ["kinetic", "external", "hartree", "exchangeCorrelation", "nuclearRepulsion", "total"]
    assert [scf["energy"][k] for k in names] == energies  # 17 digits: exact
    assert scf["energyDifference"] == -3.1e-6
    assert scf["densityDifference"] is None and scf["maxKSResidual"] is None  # NaN / inf have no JSON form
    assert scf["iterations"] == 17 and scf["converged"] is True
    mo = scf["molecularOrbitals"]
    assert [o["spin"] for o in mo] == ["alpha"] * 4 + ["beta"] * 3
    assert [o["eigenvalue"] for o in mo] == list(res["eigenvalues"]) and all(o["electrons"] == 1 for o in mo)
    assert "smearing" not in doc
    assert doc["timing"]["total"] == 0.287 and doc["counters"]["hamiltonianColumns"] == 2 ** 40 + 3


def test_result_json_smearing_section_and_buffer_size(grid, tmp_path):
    nuc, res, _ = _result(smearing=0.01)
    text = dft_amd.scf_result_json(grid, nuc, res)
    doc = json.loads(text)
    assert doc["smearing"] == {"kT": 0.01, "fermiLevel": [-0.25, -0.2], "entropyEnergy": -1e-4}
    assert all(o["electrons"] is None for o in doc["scf"]["molecularOrbitals"])
    L = dft_amd.core_lib()
    eig = res["eigenvalues"]
    n = C.c_size_t()
    args = [C.byref(grid), C.c_int(2), nuc.ctypes.data_as(C.POINTER(C.c_double)), C.byref(res["config"]),
            C.byref(res["report"]), eig.ctypes.data_as(C.POINTER(C.c_double)), C.c_int(4), C.c_int(3)]
    assert L.dft_scf_result_json(*args, None, C.c_size_t(0), C.byref(n)) == 0 and n.value == len(text)
    buf = C.create_string_buffer(n.value)  # no room for the NUL
    assert L.dft_scf_result_json(*args, buf, C.c_size_t(n.value), C.byref(n)) == -1
    # the document and the densities' field files
    res["rho_alpha"] = np.full(grid.size, 0.5)
    res["rho_beta"] = np.zeros(grid.size)
    dft_amd.save_scf_result(tmp_path / "run.json", grid, nuc, res)
    assert json.loads((tmp_path / "run.json").read_text()) == doc
    _, ra, name = dft_amd.load_field(str(tmp_path / "run.json") + ".rho_alpha.dftf")
    assert name == "rho_alpha" and np.array_equal(ra, res["rho_alpha"])
    assert math.isfinite(doc["scf"]["energy"]["total"])
