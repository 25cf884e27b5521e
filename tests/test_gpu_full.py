"""GPU parity at the benchmark's full sizes (2^28 int32, 10^8+17, 2^27 pairs
and uint64), where the CPU oracle would take too long inside a test: the
properties that fix a sort's output without computing it.

* sortedness in the element order (unsigned order of the internal bits);
* the output is a permutation of the input: exact value histograms where
  the key range is small, otherwise order-independent sums over the element
  words (plain, two non-linear mixes, squares and cross products of their
  halves) plus a 2^16-bucket histogram of the top bits;
* pairs carry value = index, so the values must be a permutation of
  0..n-1 and every output key must equal the input key at its value's
  position; with (key, value) strictly increasing this is the unique sorted
  order, exactly;
* idempotence: sorting the sorted output changes no bit.

The reference's own check is the sortedness scan (src/testMM.cu:198-213);
the bit-exact comparison with std::sort at these sizes is bench.py's."""
import pytest

pytestmark = pytest.mark.gpu

SIGN64 = -(1 << 63)


def _u32(t):
    """int32 tensor -> int64 tensor of the unsigned 32-bit words."""
    return t.to(dtype=__import__("torch").int64) & 0xFFFFFFFF


def _fmix(x, c1, c2):
    """A 32-bit xorshift-multiply finaliser (non-linear in the bits); odd
    multipliers below 2^31 keep every product below 2^63."""
    x = x ^ (x >> 16)
    x = (x * c1) & 0xFFFFFFFF
    x = x ^ (x >> 15)
    x = (x * c2) & 0xFFFFFFFF
    return x ^ (x >> 16)


def _hash_sums(words):
    """Order-independent sums over int64 words < 2^32: the words, two
    finaliser mixes, and the squares and cross products of their 16-bit
    halves (every term below 2^32, every sum of up to 2^28 terms below 2^61)."""
    lo, hi = words & 0xFFFF, words >> 16
    terms = (words, _fmix(words, 0x7FEB352D, 0x446CA68B), _fmix(words, 0x2C1B3C6D, 0x297A2D39), lo * lo, lo * hi)
    return [int(t.sum()) for t in terms]


def _multiset_u32(words):
    torch = __import__("torch")
    return _hash_sums(words), torch.bincount(words >> 16, minlength=1 << 16)


def _same_multiset_u32(x, y):
    torch = __import__("torch")
    hx, bx = _multiset_u32(x)
    hy, by = _multiset_u32(y)
    return hx == hy and torch.equal(bx, by)


def _nondecreasing(keys):
    return bool((keys[1:] >= keys[:-1]).all())


def _gen(seed, dev):
    torch = __import__("torch")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    return g


def _int32_case(gmms, dev, n, lo, hi, seed, exact_hist):
    torch = __import__("torch")
    print(f"full int32 n={n} range=[{lo},{hi})", flush=True)
    x = torch.randint(lo, hi, (n,), dtype=torch.int32, device=dev, generator=_gen(seed, dev))
    y = gmms.sort(x)
    assert y.shape == x.shape
    assert _nondecreasing(y)  # int32 compares signed, as the element order does
    if exact_hist:
        assert torch.equal(torch.bincount((x - lo).to(torch.int64), minlength=hi - lo),
                           torch.bincount((y - lo).to(torch.int64), minlength=hi - lo))
    else:
        assert _same_multiset_u32(_u32(x), _u32(y))
    z = gmms.sort(y)
    assert torch.equal(z, y)


def test_int32_2p28_full_range(gmms, dev):
    _int32_case(gmms, dev, 1 << 28, -(1 << 31), (1 << 31) - 1, 0x6D6D7301, False)


def test_int32_1e8_plus_17(gmms, dev):
    _int32_case(gmms, dev, 10**8 + 17, 0, 1 << 30, 0x6D6D7302, False)


def test_int32_2p28_duplicate_heavy(gmms, dev):
    _int32_case(gmms, dev, 1 << 28, 0, 1024, 0x6D6D7303, True)


def test_float32_2p28_total_order(gmms, dev):
    torch = __import__("torch")
    n = 1 << 28
    print(f"full float32 n={n}", flush=True)
    x = torch.randn((n,), dtype=torch.float32, device=dev, generator=_gen(0x6D6D7304, dev))
    x[:4] = torch.tensor([0.0, -0.0, float("inf"), -float("inf")], device=dev)
    y = gmms.sort(x)
    u = _u32(y.view(torch.int32))
    key = torch.where(u >= (1 << 31), u ^ 0xFFFFFFFF, u | (1 << 31))  # IEEE total order
    assert _nondecreasing(key)
    assert _same_multiset_u32(_u32(x.view(torch.int32)), u)
    assert torch.equal(gmms.sort(y).view(torch.int32), y.view(torch.int32))


def _u64(dev, n, seed):
    torch = __import__("torch")
    g = _gen(seed, dev)
    hi = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32, device=dev, generator=g)
    lo = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32, device=dev, generator=g)
    return (hi.to(torch.int64) << 32) | _u32(lo)


@pytest.mark.parametrize("dup_shift", [0, 44])
def test_pairs_2p27_value_is_index(gmms, dev, dup_shift):
    """(uint64 key, uint32 value = index); dup_shift 44 leaves 2^20 distinct
    keys, so most ties are broken by the value."""
    torch = __import__("torch")
    n = 1 << 27
    print(f"full pairs n={n} dup_shift={dup_shift}", flush=True)
    k = _u64(dev, n, 0x6D6D7310 + dup_shift)
    if dup_shift:
        k = (k >> dup_shift) & ((1 << (64 - dup_shift)) - 1)
    v = torch.arange(n, dtype=torch.int32, device=dev)
    ks, vs = gmms.sort(k, v)
    idx = _u32(vs)
    assert torch.equal(torch.bincount(idx, minlength=n), torch.ones(n, dtype=torch.int64, device=dev))
    assert torch.equal(ks, k[idx])
    s = ks ^ SIGN64  # unsigned order as signed
    strictly = (s[:-1] < s[1:]) | ((s[:-1] == s[1:]) & (idx[:-1] < idx[1:]))
    assert bool(strictly.all())


def test_uint64_2p27(gmms, dev):
    torch = __import__("torch")
    n = 1 << 27
    print(f"full uint64 n={n}", flush=True)
    k = _u64(dev, n, 0x6D6D7320)

    def sort_u64(x):
        a, b = x.clone(), torch.empty_like(x)
        return b if gmms.sort_inplace(a, b, dtype=gmms.UINT64) else a

    y = sort_u64(k)
    assert _nondecreasing(y ^ SIGN64)  # unsigned order as signed
    lo = lambda t: t & 0xFFFFFFFF
    hi = lambda t: (t >> 32) & 0xFFFFFFFF
    mix = lambda t: lo(t) ^ ((hi(t) * 0x5BD1E995) & 0xFFFFFFFF)  # one 32-bit word per element
    assert _same_multiset_u32(lo(k), lo(y)) and _same_multiset_u32(hi(k), hi(y))
    assert _same_multiset_u32(mix(k), mix(y))
    assert torch.equal(sort_u64(y), y)
