"""GPU test of fftgal_qpm_load_device (include/fftgal_batch.h): a QPM
catalog file parsed on host threads into pinned staging buffers, a piece at a
time, with each piece's copy to the device overlapping the next piece's
parse.  The device arrays must equal the oracle's stdio restatement of the
reference loader (src/io/qpm.c:9-34) rounded to float32, bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _write_qpm(path, n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * 2000.0
    vel = rng.standard_normal((n, 3)) * 300.0
    with open(path, "w") as fp:
        for p, v in zip(pos, vel):
            fp.write(f"{p[0]:.6f} {p[1]:.6f} {p[2]:.6f} {v[0]:.3f} {v[1]:.3f} {v[2]:.3f} "
                     f"1.234e+12 0 {seed}\n")


@pytest.mark.parametrize("chunk,piece", [(None, None), ("4096", "1024")])
def test_qpm_load_device_matches_oracle(torch_cuda, lib, orc, tmp_path, monkeypatch, chunk,
                                        piece):
    torch = torch_cuda
    if chunk:   # many chunks and pieces: both staging buffers cycle
        monkeypatch.setenv("FFTGAL_QPM_CHUNK_BYTES", chunk)
        monkeypatch.setenv("FFTGAL_QPM_PIECE_LINES", piece)
    n = 30000
    path = str(tmp_path / "cat.qpm")
    _write_qpm(path, n, 4)
    o = orc.load().oracle_loadqpm_cubic(path.encode(), 2000.0)
    ref = [np.ctypeslib.as_array(getattr(o.contents, k), shape=(n,)).astype(np.float32)
           for k in ("x", "y", "z")]
    orc.load().oracle_gal_free(o)
    dev = [torch.full((n + 7,), -1.0, dtype=torch.float32, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    got_n = C.c_long()
    rc = lib.fftgal_qpm_load_device(path.encode(), n + 7, *(d.data_ptr() for d in dev),
                                    C.byref(got_n))
    assert rc == 0 and got_n.value == n
    for d, r in zip(dev, ref):
        h = d.cpu().numpy()
        assert np.array_equal(h[:n], r)
        assert np.all(h[n:] == -1.0)      # nothing written past the catalog


def test_qpm_load_device_errors(torch_cuda, lib, tmp_path):
    torch = torch_cuda
    path = str(tmp_path / "c.qpm")
    _write_qpm(path, 100, 5)
    d = [torch.zeros(64, device="cuda") for _ in range(3)]
    n = C.c_long(-1)
    assert lib.fftgal_qpm_load_device(path.encode(), 64, *(t.data_ptr() for t in d),
                                      C.byref(n)) == -2          # more lines than Np_max
    assert n.value == 0
    assert lib.fftgal_qpm_load_device(str(tmp_path / "missing.qpm").encode(), 64,
                                      *(t.data_ptr() for t in d), C.byref(n)) < 0
    bad = tmp_path / "bad.qpm"
    bad.write_text("1 2 3 4 5 6 7 0 1\n1 2 x 4 5 6 7 0 1\n")
    assert lib.fftgal_qpm_load_device(str(bad).encode(), 64, *(t.data_ptr() for t in d),
                                      C.byref(n)) == -2
