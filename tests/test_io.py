"""Gauge-configuration text files (src/QED2dLoadSave.jl:1-19): writedlm of
hcat(real, imag) of reshape(U, 2V) -- one link per line, column-major order.
Host-only entry points; no GPU needed."""
import numpy as np

from dense import random_links


def test_roundtrip_bit_exact(tmp_path):
    import qed2d
    rng = np.random.default_rng(0)
    U = random_links(7, 5, rng)
    U[0, 0, 0] = 1.0
    U[1, 2, 3] = -0.0 + 1e-300j
    f = tmp_path / "cfg"
    qed2d.write_links_text(U, f)
    V = qed2d.read_links_text(f, 7, 5)
    assert np.array_equal(U.view(np.float64), V.view(np.float64))


def test_layout_matches_julia_writedlm(tmp_path):
    """Line k holds link k of the column-major [L1, L2, 2] array: i1 fastest."""
    import qed2d
    L1, L2 = 4, 3
    U = np.zeros((2, L2, L1), complex)
    for mu in range(2):
        for i2 in range(L2):
            for i1 in range(L1):
                U[mu, i2, i1] = (i1 + 10 * i2 + 100 * mu) + 0.5j
    f = tmp_path / "cfg"
    qed2d.write_links_text(U, f)
    rows = np.loadtxt(f, delimiter="\t")
    assert rows.shape == (2 * L1 * L2, 2)
    k = np.arange(2 * L1 * L2)
    np.testing.assert_array_equal(rows[:, 0], (k % L1) + 10 * ((k // L1) % L2) + 100 * (k // (L1 * L2)))
    np.testing.assert_array_equal(rows[:, 1], 0.5)
    # a file written by numpy in the same layout (as Julia's writedlm would) reads back
    g = tmp_path / "cfg2"
    np.savetxt(g, np.column_stack([U.real.ravel(), U.imag.ravel()]), delimiter="\t", fmt="%.17g")
    assert np.array_equal(qed2d.read_links_text(g, L1, L2), U)


def test_short_or_long_file_rejected(tmp_path):
    import pytest
    import qed2d
    f = tmp_path / "cfg"
    qed2d.write_links_text(np.ones((2, 3, 3), complex), f)
    with pytest.raises(qed2d.QED2dError):
        qed2d.read_links_text(f, 4, 3)
    with pytest.raises(qed2d.QED2dError):
        qed2d.read_links_text(f, 2, 3)


def test_read_from_fifo(tmp_path):
    """A path that cannot be seeked (a FIFO) reads like a regular file."""
    import os
    import threading

    import qed2d
    rng = np.random.default_rng(3)
    U = random_links(6, 4, rng)
    f = tmp_path / "cfg"
    qed2d.write_links_text(U, f)
    data = f.read_bytes()
    fifo = tmp_path / "fifo"
    os.mkfifo(fifo)

    def writer():
        with open(fifo, "wb") as w:
            w.write(data)

    th = threading.Thread(target=writer, daemon=True)
    th.start()
    try:
        V = qed2d.read_links_text(fifo, 6, 4)
    finally:
        th.join(timeout=10)
    assert np.array_equal(U.view(np.float64), V.view(np.float64))
