"""The allocation-free radix sort (mpm_dev_radix_sort_noalloc) against the
entry point with a caller-allocated temp buffer (mpm_dev_radix_sort), through
the C ABI: two calls in a row on one device, the second with a smaller input,
each bit-exact against the existing entry point and a stable CPU sort."""
import numpy as np
import pytest

from squishy_volumes_amd import abi

pytestmark = pytest.mark.gpu


def sort_with_temp(lib, keys, values, bits):
    import torch

    n = keys.size
    k_in = torch.from_numpy(keys).cuda()
    v_in = torch.from_numpy(values).cuda()
    k_out = torch.empty_like(k_in)
    v_out = torch.empty_like(v_in)
    temp = torch.empty(max(int(lib.mpm_dev_radix_sort_temp_bytes(n)), 16), dtype=torch.uint8, device="cuda")
    assert lib.mpm_dev_radix_sort(abi.ptr(k_in), abi.ptr(v_in), abi.ptr(k_out), abi.ptr(v_out), n, bits,
                                  abi.ptr(temp), None) == abi.OK
    torch.cuda.synchronize()
    return k_out.cpu().numpy(), v_out.cpu().numpy()


def sort_without_temp(lib, keys, values, bits):
    import torch

    n = keys.size
    k_in = torch.from_numpy(keys).cuda()
    v_in = torch.from_numpy(values).cuda()
    k_out = torch.empty_like(k_in)
    v_out = torch.empty_like(v_in)
    torch.cuda.synchronize()
    assert lib.mpm_dev_radix_sort_noalloc(abi.ptr(k_in), abi.ptr(v_in), abi.ptr(k_out), abi.ptr(v_out), n, bits,
                                          None) == abi.OK
    torch.cuda.synchronize()
    return k_out.cpu().numpy(), v_out.cpu().numpy()


def test_noalloc_sort_matches_the_temp_buffer_sort_on_two_calls(engine_lib):
    lib = engine_lib
    cap = int(lib.mpm_dev_radix_sort_noalloc_max_keys())
    assert cap == 1 << 22
    rng = np.random.default_rng(11)
    # the largest input the array takes, then a smaller one on the same array
    for n, bits in ((cap, 18), (100_003, 24)):
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint32)
        keys[: n // 3] = keys[0]  # a long run of equal keys: stability shows
        values = rng.permutation(n).astype(np.uint32)
        k0, v0 = sort_with_temp(lib, keys, values, bits)
        k1, v1 = sort_without_temp(lib, keys, values, bits)
        assert np.array_equal(k1, k0) and np.array_equal(v1, v0)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k1, keys[order]) and np.array_equal(v1, values[order])


def test_noalloc_sort_refuses_inputs_beyond_its_array(engine_lib):
    lib = engine_lib
    cap = int(lib.mpm_dev_radix_sort_noalloc_max_keys())
    # refused before any device access: the (host) pointers are never dereferenced
    bufs = [np.zeros(4, np.uint32) for _ in range(4)]
    k_in, v_in, k_out, v_out = (abi.ptr(b) for b in bufs)
    assert lib.mpm_dev_radix_sort_noalloc(k_in, v_in, k_out, v_out, cap + 1, 24, None) == abi.ERR_CAPACITY
