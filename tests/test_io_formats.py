"""The reference's on-disk formats through the C-ABI (no GPU needed for host
arrays): the density/potential file of write_data / read_data
(io_module.f90:68-455) and the wavefunction file of simple_wf_io_write /
simple_wf_io_read (io_write_module.f90:19-264, io_read_module.f90:15-521).

Pinned by: the file the reference itself wrote in its Si2 SCF example
(tests/golden/si2_vrho.dat1: the density integrates to the 8 valence
electrons of Si2); an independent record-by-record parse of what the library
writes; a file written with a shuffled grid map; and a Fortran program
(flang) that reads the wavefunction file with the reference's read sequence.
"""
import json
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from rsdft_amd import gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def records(path):
    """Fortran sequential unformatted records (4-byte length frames)."""
    data = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(data):
        (n,) = struct.unpack("<i", data[pos:pos + 4])
        (m,) = struct.unpack("<i", data[pos + 4 + n:pos + 8 + n])
        assert n == m
        out.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def write_records(path, recs):
    with open(path, "wb") as f:
        for r in recs:
            f.write(struct.pack("<i", len(r)) + r + struct.pack("<i", len(r)))


def test_reads_the_reference_si2_density_file():
    meta = json.load(open(os.path.join(GOLD, "si2_vrho.json")))
    N = tuple(meta["ngrid"])
    ML = N[0] * N[1] * N[2]
    rho, vloc, vh, vxc, ef = gpu.read_vrho(os.path.join(GOLD, "si2_vrho.dat1"), N)
    aa = meta["ax"] * np.array([meta["a1"], meta["a2"], meta["a3"]])
    dV = abs(np.linalg.det(aa)) / ML
    assert abs(rho.sum() * dV - meta["valence_electrons"]) < 1e-8
    assert ef == 0.0                                   # unversioned file
    # the same values as an independent parse (the file's LL2 is the natural order)
    rec = records(os.path.join(GOLD, "si2_vrho.dat1"))
    assert np.frombuffer(rec[0], "<i4").tolist() == [ML, *N]
    LL = np.frombuffer(rec[1], "<i4").reshape(ML, 3)
    k = LL[:, 0] + N[0] * (LL[:, 1] + N[1] * LL[:, 2])
    for j, arr in enumerate((rho[0], vloc[0], vh, vxc[0])):
        want = np.empty(ML)
        want[k] = np.frombuffer(rec[2 + j], "<f8")
        assert np.array_equal(arr, want)
    # the periodic Hartree potential has no G = 0 component
    assert abs(np.mean(vh)) < 1e-12


def test_vrho_roundtrip_versioned(tmp_path):
    N = (6, 8, 10)
    ML = 480
    rng = np.random.default_rng(1)
    rho, vloc, vxc = (rng.standard_normal((2, ML)) for _ in range(3))
    vh = rng.standard_normal(ML)
    aa = np.diag([3.0, 4.0, 5.0])
    at = rng.uniform(0, 1, (2, 3))
    p = str(tmp_path / "vrho.dat")
    gpu.write_vrho(p, N, rho, vloc, vh, vxc, aa=aa, aa_atom=at, ki_atom=[1, 1], zn_atom=[14],
                   efermi=0.25)
    r2, l2, h2, x2, ef = gpu.read_vrho(p, N, MSP=2)
    assert ef == 0.25
    for a, b in ((rho, r2), (vloc, l2), (vh, h2), (vxc, x2)):
        assert np.array_equal(a, b)
    rec = records(p)
    assert rec[0].decode().rstrip() == "version3.0, comment_length=64" and len(rec[0]) == 64
    assert np.frombuffer(rec[1], "<i4").tolist() == [ML, *N]
    assert np.array_equal(np.frombuffer(rec[4], "<f8"), aa.T.ravel())   # column-major aa
    assert np.frombuffer(rec[6], "<i4").tolist() == [1, 2]
    assert struct.unpack("<d", rec[11])[0] == 0.25
    # spin 1: rho, Vloc, Vh, Vxc; spin 2: rho, Vloc, Vxc
    assert len(rec) == 12 + 4 + 3
    assert np.array_equal(np.frombuffer(rec[14], "<f8"), vh)


@pytest.mark.parametrize("OC", [1, 4, 11])
def test_wf_roundtrip_and_layout(tmp_path, OC):
    N = (6, 8, 10)
    ML, MB = 480, 7
    rng = np.random.default_rng(OC)
    unk = rng.standard_normal((MB, ML))
    occ = np.array([2.0] * 4 + [0.0] * 3)
    p = str(tmp_path / f"wf{OC}")
    gpu.write_wf(p, N, unk, occ, OC=OC, aa=np.eye(3) * 5.0, bb=np.eye(3) * 0.2)
    back, occ2, mbf = gpu.read_wf(p, N, MB)
    assert mbf == MB and np.array_equal(occ2, occ)
    want = unk.astype(np.float32).astype(np.float64) if OC == 4 else unk
    assert np.array_equal(back, want)
    rec = records(p)
    if OC < 10:
        assert struct.unpack("<i", rec[0])[0] == -1
        assert np.frombuffer(rec[1], "<i4").tolist() == [ML, *N]
        assert np.frombuffer(rec[2], "<i4").tolist() == [MB, 1, MB]
        assert np.frombuffer(rec[3], "<i4").tolist() == [1, 1, 1]
        assert np.frombuffer(rec[4], "<i4").tolist() == [0, OC, 1]
        LL = np.frombuffer(rec[5], "<i4").reshape(ML, 3)
        assert np.frombuffer(rec[6], "<f8").tolist() == occ.tolist()
        assert len(rec[7]) == 21 * 8 and np.frombuffer(rec[8], "<i4").tolist() == [1] * 8
        assert [np.frombuffer(r, "<i4").tolist() for r in rec[9:13]] == [[0, ML], [0, MB],
                                                                         [0, 1], [0, 1]]
        bands = rec[13:]
    else:
        assert np.frombuffer(rec[0], "<i4").tolist() == [ML, *N]
        LL = np.frombuffer(rec[2], "<i4").reshape(ML, 3)
        bands = rec[4:]
    assert LL[:3].tolist() == [[0, 0, 0], [1, 0, 0], [2, 0, 0]] and LL[-1].tolist() == [5, 7, 9]
    assert len(bands) == MB
    dt = "<f4" if OC == 4 else "<f8"
    assert np.array_equal(np.frombuffer(bands[3], dt).astype(np.float64), want[3])


def test_wf_band_range_and_shuffled_grid_map(tmp_path):
    """A file written by another decomposition: its LL2 lists the points in a
    different order and it holds bands 2..5 of 6; the reader maps each value
    through LL2 and leaves the other bands alone."""
    N = (4, 4, 8)
    ML, MBf = 128, 6
    rng = np.random.default_rng(3)
    perm = rng.permutation(ML)
    i1, i2, i3 = perm % 4, (perm // 4) % 4, perm // 16
    LL = np.stack([i1, i2, i3], axis=1).astype("<i4")
    vals = rng.standard_normal((MBf, ML))
    recs = [struct.pack("<i", -1), np.array([ML, *N], "<i4").tobytes(),
            np.array([MBf, 2, 5], "<i4").tobytes(), np.array([1, 1, 1], "<i4").tobytes(),
            np.array([0, 1, 1], "<i4").tobytes(), LL.tobytes(),
            np.arange(MBf, dtype="<f8").tobytes(), np.zeros(21).tobytes(),
            np.ones(8, "<i4").tobytes()] + [np.array([0, 1], "<i4").tobytes()] * 4
    recs += [vals[n][perm].astype("<f8").tobytes() for n in range(1, 5)]
    p = str(tmp_path / "wf_shuffled")
    write_records(p, recs)
    unk = np.full((4, ML), 7.0)
    back, occ, mbf = gpu.read_wf(p, N, 4, unk=unk)
    assert mbf == MBf and occ.tolist() == [0.0, 1.0, 2.0, 3.0]
    assert np.all(back[0] == 7.0)                       # band 1 is not in the file
    assert np.array_equal(back[1:4], vals[1:4])         # bands 2..4, file band 5 beyond MB


def test_wrong_grid_is_refused(tmp_path):
    p = str(tmp_path / "wf")
    gpu.write_wf(p, (4, 4, 4), np.zeros((2, 64)), np.zeros(2))
    with pytest.raises(gpu.RsdftGpuError, match="grid"):
        gpu.read_wf(p, (4, 4, 8), 2)


FREAD = r"""
program read_wf_check
  ! the read sequence of simple_wf_io_read (io_read_module.f90:131-325) for a
  ! new-format real(8) file of one process; prints ML, MB and the sum of
  ! (band index x value) over the bands read
  implicit none
  integer :: itmp(21), ML, MB, MB1, MB2, n
  integer, allocatable :: LL(:,:)
  real(8), allocatable :: occ(:), dtmp(:)
  real(8) :: aa(3,3), bb(3,3), kbb(3,1), s
  character(len=1024) :: f
  call get_command_argument(1, f)
  open(3, file=trim(f), form='unformatted', status='old')
  read(3) itmp(1)
  read(3) itmp(1:4)
  read(3) itmp(5:7)
  read(3) itmp(8:10)
  read(3) itmp(11:13)
  ML = itmp(1); MB = itmp(5); MB1 = itmp(6); MB2 = itmp(7)
  allocate(LL(3, ML), occ(MB), dtmp(ML))
  read(3) LL
  read(3) occ
  read(3) aa, bb, kbb
  read(3) itmp(14:21)
  read(3)
  read(3)
  read(3)
  read(3)
  s = 0.0d0
  do n = MB1, MB2
    read(3) dtmp
    s = s + n * sum(dtmp)
  end do
  close(3)
  write(*, '(2i8, es25.16)') ML, MB, s
end program read_wf_check
"""


def test_fortran_reads_the_wf_file(tmp_path):
    flang = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
    if not os.path.exists(flang):
        pytest.skip("no Fortran compiler")
    src = tmp_path / "read_wf_check.f90"
    src.write_text(FREAD)
    exe = str(tmp_path / "read_wf_check")
    subprocess.run([flang, "-O1", str(src), "-o", exe], check=True, cwd=tmp_path)
    N = (6, 8, 10)
    rng = np.random.default_rng(9)
    unk = rng.standard_normal((5, 480))
    p = str(tmp_path / "wf")
    gpu.write_wf(p, N, unk, np.ones(5), OC=1, aa=np.eye(3), bb=np.eye(3))
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.dirname(flang) + "/../lib:" +
               os.environ.get("LD_LIBRARY_PATH", ""))
    out = subprocess.run([exe, p], check=True, capture_output=True, text=True, env=env).stdout
    ML, MB, s = out.split()
    assert int(ML) == 480 and int(MB) == 5
    want = sum((n + 1) * unk[n].sum() for n in range(5))
    assert abs(float(s) - want) < 1e-12 * max(1.0, abs(want))
