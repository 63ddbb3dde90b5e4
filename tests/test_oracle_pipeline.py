"""CPU checks of the gradient / structure-tensor oracle (volume.h:974-1021):
linear fields are differentiated exactly everywhere, quadratic fields
exactly inside and by the one-sided values f' -/+ f''/2 at the ends."""
import numpy as np

from oracle import oracle as O


def _grid(shape):
    return np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")


def test_linear_field_exact_everywhere():
    z, y, x = _grid((5, 6, 7))
    for sp in ((1.0, 1.0, 1.0), (0.5, 2.0, 4.0)):
        f = (3 * x - 2 * y + 5 * z).astype(np.float32)
        gx, gy, gz = O.gradient3(f, sp)
        assert np.all(gx == np.float32(3 / sp[0])) and np.all(gy == np.float32(-2 / sp[1]))
        assert np.all(gz == np.float32(5 / sp[2]))


def test_quadratic_field_interior_and_ends():
    z, y, x = _grid((4, 5, 9))
    f = (x * x).astype(np.float32)
    gx, gy, gz = O.gradient3(f)
    assert np.all(gx[..., 1:-1] == 2 * x[..., 1:-1])
    assert np.all(gx[..., 0] == 1.0)            # forward difference: f'(0) + f''/2
    assert np.all(gx[..., -1] == 2 * 8 - 1)     # backward difference: f'(8) - f''/2
    assert not gy.any() and not gz.any()


def test_structure_tensor_is_outer_product():
    rng = np.random.default_rng(0)
    f = rng.normal(size=(6, 7, 8)).astype(np.float32)
    t = O.structure_tensor3(f, (1.0, 1.5, 0.75))
    gx, gy, gz = O.gradient3(f, (1.0, 1.5, 0.75))
    assert np.array_equal(t[..., 1], gx * gy) and np.array_equal(t[..., 5], gz * gz)
    M = np.stack([np.stack([t[..., 0], t[..., 1], t[..., 2]], -1), np.stack([t[..., 1], t[..., 3], t[..., 4]], -1),
                  np.stack([t[..., 2], t[..., 4], t[..., 5]], -1)], -2).astype(np.float64)
    ev = np.linalg.eigvalsh(M)
    assert np.allclose(ev[..., :2], 0, atol=1e-5 * np.abs(ev).max())  # rank one
