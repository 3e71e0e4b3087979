"""GPU: the allocation-free weight gradient (shy_nn_conv_wgrad_noalloc)
against the workspace entry point (shy_nn_conv_wgrad) it stands beside.

Both are called through the C ABI on the same bf16 inputs; the workspace
call gets the split count the allocation-free one derives from its fixed
scratch (min(SHY_NN_NOALLOC_FLOATS / (Cout*taps*Cin), SHY_NN_NOALLOC_MAX_SPLITS)),
so the two sum the same partials in the same order and must agree bit for
bit.  Two calls in a row, the second on a smaller input, share the scratch
array: the second must not see the first's partials.  A float32 torch
reference on the CPU checks the values themselves.
"""
import os
import re

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_constant(name):
    text = open(os.path.join(ROOT, "include", "shy_nn.h")).read()
    return int(re.search(r"#define\s+" + name + r"\s+(\d+)", text).group(1))


@pytest.fixture(scope="module")
def sg(shy):
    from simplehydrology_amd import surrogate
    surrogate.nn()
    return surrogate


def ref_wgrad(x, g, taps, cin, cout):
    """dW[co][tap*Cin + ci] = sum_p g[p][co] x[p + tap][ci] in float32."""
    k = 3 if taps == 9 else 1
    xr = x.clone().requires_grad_(False)
    w = torch.zeros(cout, cin, k, k, requires_grad=True)
    y = torch.nn.functional.conv2d(xr.permute(0, 3, 1, 2), w, None, padding=k // 2)
    y.backward(g.permute(0, 3, 1, 2))
    return w.grad.permute(0, 2, 3, 1).reshape(cout, taps * cin)


@pytest.mark.parametrize("shapes", [
    [(2, 64, 48, 128, 128, 9), (1, 16, 24, 32, 64, 9)],   # LDS kernel path
    [(2, 40, 24, 8, 32, 9), (1, 13, 7, 16, 8, 1)],        # direct-load path
])
def test_noalloc_wgrad_equals_workspace_wgrad(sg, shapes):
    L = sg.nn()
    cap = header_constant("SHY_NN_NOALLOC_FLOATS")
    max_splits = header_constant("SHY_NN_NOALLOC_MAX_SPLITS")
    stream = torch.cuda.current_stream().cuda_stream
    for call, (N, H, W, cin, cout, taps) in enumerate(shapes):
        print("noalloc wgrad call", call, (N, H, W, cin, cout, taps), flush=True)
        gen = torch.Generator().manual_seed(100 + call)
        x = torch.randn(N, H, W, cin, generator=gen).to(torch.bfloat16)
        g = torch.randn(N, H, W, cout, generator=gen).to(torch.bfloat16)
        xd, gd = x.cuda().contiguous(), g.cuda().contiguous()
        per = cout * taps * cin
        splits = min(cap // per, max_splits)
        part = torch.empty(splits * per, dtype=torch.float32, device="cuda")
        want = torch.empty(cout, taps * cin, dtype=torch.float32, device="cuda")
        got = torch.full((cout, taps * cin), float("nan"), dtype=torch.float32, device="cuda")
        rc = L.shy_nn_conv_wgrad(xd.data_ptr(), gd.data_ptr(), want.data_ptr(), part.data_ptr(),
                                 splits, N, H, W, cin, cout, taps, stream)
        assert rc == 0, L.shy_nn_last_error()
        rc = L.shy_nn_conv_wgrad_noalloc(xd.data_ptr(), gd.data_ptr(), got.data_ptr(), N, H, W,
                                         cin, cout, taps, stream)
        assert rc == 0, L.shy_nn_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.int32), want.view(torch.int32)), call
        ref = ref_wgrad(x.float(), g.float(), taps, cin, cout)
        err = float((got.cpu().double() - ref.double()).norm() / ref.double().norm())
        assert err < 1e-4, (call, err)  # FP32 accumulation of bf16 products


def test_noalloc_wgrad_rejects_oversized_layers(sg):
    L = sg.nn()
    stream = torch.cuda.current_stream().cuda_stream
    x = torch.zeros(8, dtype=torch.bfloat16, device="cuda")
    dw = torch.zeros(8, dtype=torch.float32, device="cuda")
    # 1024 x 9 x 512 floats of partials per split exceed the 4 Mi-float scratch
    rc = L.shy_nn_conv_wgrad_noalloc(x.data_ptr(), x.data_ptr(), dw.data_ptr(), 1, 1, 1, 512,
                                     1024, 9, stream)
    assert rc != 0
    assert "exceeds the scratch" in L.shy_nn_last_error().decode()
