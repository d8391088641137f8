"""MatrixMarket readers of the solve-only mode (read_mm_coordinate_real,
read_mm_array_vector_mx1, to_upper_triangle_block_triplets, main.cu:387-590),
host code of libsgipc.so, pinned by scipy.io's independent MatrixMarket
implementation: vectors round-trip exactly; a general file yields its upper
3x3 blocks (lower ones dropped), a symmetric one the mirrored blocks; bad
shapes and missing diagonal blocks are refused."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp


def _spd_blocks(nb, rng):
    A = sp.random(3 * nb, 3 * nb, density=0.2, random_state=rng, format="csr")
    A = A + A.T + sp.identity(3 * nb) * 10
    return A.toarray()


def test_vector_round_trip(tmp_path):
    import sgipc
    rng = np.random.default_rng(0)
    b = rng.normal(size=(30, 1)) * 10.0 ** rng.integers(-8, 8, size=(30, 1))
    path = tmp_path / "b.mtx"
    scipy.io.mmwrite(str(path), b, precision=17)
    got = sgipc.read_mm_vector(path)
    assert np.array_equal(got, b[:, 0])


@pytest.mark.parametrize("symmetry", ["general", "symmetric"])
def test_upper_blocks_of_a_coordinate_file(tmp_path, symmetry):
    import sgipc
    rng = np.random.default_rng(1)
    D = _spd_blocks(7, rng)
    path = tmp_path / "A.mtx"
    scipy.io.mmwrite(str(path), sp.coo_matrix(D), symmetry=symmetry, precision=17)
    nb, r, c, v = sgipc.read_mm_upper_bcoo(path)
    assert nb == 7 and np.all(r <= c)
    keys = list(zip(r.tolist(), c.tolist()))
    assert keys == sorted(keys)
    U = np.zeros_like(D)
    for (i, j), blk in zip(keys, v):
        U[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
    # the upper block triangle of D, diagonal blocks whole
    want = np.zeros_like(D)
    for i in range(7):
        for j in range(i, 7):
            want[3 * i:3 * i + 3, 3 * j:3 * j + 3] = D[3 * i:3 * i + 3, 3 * j:3 * j + 3]
    assert np.array_equal(U, want)


def test_refusals(tmp_path):
    import sgipc
    p = tmp_path / "bad.mtx"
    scipy.io.mmwrite(str(p), sp.coo_matrix(np.eye(4)), precision=17)
    with pytest.raises(sgipc.SgipcError, match="divisible by 3"):
        sgipc.read_mm_upper_bcoo(p)
    D = np.eye(6)
    D[3:, 3:] = 0.0   # block 1 has no diagonal block
    D[0, 4] = D[4, 0] = 1.0
    scipy.io.mmwrite(str(p), sp.coo_matrix(D), precision=17)
    with pytest.raises(sgipc.SgipcError, match="missing diagonal"):
        sgipc.read_mm_upper_bcoo(p)
    scipy.io.mmwrite(str(p), np.ones((3, 2)), precision=17)
    with pytest.raises(sgipc.SgipcError, match="M x 1"):
        sgipc.read_mm_vector(p)
