"""The whole-array lozenge face cover used by the legality checks equals the face-by-face form,
on tilings of almost-hexagons (reference constructors) and on arbitrary 6-bit state arrays."""
import numpy as np
import pytest

from randomtilings_amd import statics as S
from tests.tilings import lozenge_face_cover, lozenge_face_cover_loop


@pytest.mark.parametrize("a", [5, 10, 20])
@pytest.mark.parametrize("which", ["MinTiling", "MaxTiling"])
def test_cover_on_tilings(a, which):
    t = S.call(f"lozenge.{which}", S.call("lozenge.AlmostHexDomain", args=[a, 0]))
    assert np.array_equal(lozenge_face_cover(t), lozenge_face_cover_loop(t))


def test_cover_on_arbitrary_states():
    rng = np.random.default_rng(1)
    for n in (4, 13, 37):
        t = rng.integers(0, 64, n * n).astype(np.int32)
        assert np.array_equal(lozenge_face_cover(t), lozenge_face_cover_loop(t))
