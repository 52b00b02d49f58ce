"""bitset (bitset.cuh) against a numpy model of the reference semantics: per-bit set / reset / flip
return the previous bit, whole-set set() / reset() / flip(), count() over the bits below size()
(the last block's spare bits never count), all / any / none."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402


def dev(a):
    return torch.from_numpy(np.array(a, dtype=np.int64, copy=True)).cuda()


@pytest.mark.parametrize("nbits", [1, 31, 32, 33, 1000, (1 << 20) + 5])
def test_bitset_matches_model(nbits):
    rng = np.random.default_rng(nbits)
    b = sa.bitset.createDeviceObject(nbits)
    model = np.zeros(nbits, dtype=np.uint8)
    assert b.size() == nbits and b.count() == 0 and b.none() and not b.any()
    for op in ("set", "flip", "reset", "flip", "set"):
        pos = rng.permutation(nbits)[: max(1, nbits // 3)]  # unique positions: old values are exact
        old = b.update(op, dev(pos)).cpu().numpy()
        np.testing.assert_array_equal(old, model[pos])
        if op == "set":
            model[pos] = 1
        elif op == "reset":
            model[pos] = 0
        else:
            model[pos] ^= 1
        assert b.count() == int(model.sum())
    probe = rng.integers(0, nbits, size=min(nbits, 5000))
    np.testing.assert_array_equal(b.test(dev(probe)).cpu().numpy(), model[probe])
    # out-of-range positions are skipped
    before = b.count()
    assert b.update("set", dev([-1, nbits, nbits + 100])).cpu().numpy().tolist() == [0, 0, 0]
    assert b.count() == before
    # the same position twice in one call: two flips cancel, the old values are {b, not b}
    p = int(rng.integers(0, nbits))
    was = int(model[p])
    old2 = sorted(b.update("flip", dev([p, p])).cpu().numpy().tolist())
    assert old2 == sorted([was, 1 - was]) and int(b.test(dev([p])).cpu()[0]) == was
    # whole-set operations; the last block's spare bits never count
    b.set()
    assert b.count() == nbits and b.all()
    b.flip()
    assert b.count() == 0 and b.none()
    b.flip()
    assert b.all()
    b.reset()
    assert b.none()
    b.destroy()
