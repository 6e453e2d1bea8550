"""Adjoints on the GPU: the fused forward + backward chunk kernel and the product adjoint.

Bars (the adjoint has no reference counterpart; the oracle's adjoint is pinned on the CPU by
central differences, tests/test_adjoint.py):
 - every code, fp32 and fp64, interpreter and specialised kernel: gradients within
   1e-5 |g| + 1e-6 max|g| (fp32) or 1e-11 |g| + 1e-12 max|g| (fp64) of the oracle's, which
   evaluates the derivatives in double at the same forward values; the loss within 1e-6 (fp32) or
   1e-12 (fp64) of the sum of the terms' magnitudes;
 - the differentiation workload's loss sum(sinh(a*b + c) * exp(-(d*d))) on seeded uniform[-2, 2]
   fp32 data: against the oracle as above, against torch's float64 autograd of the same expression
   (driven through a torch.autograd.Function that calls the library), and against float64 central
   differences (step 1e-3) at 64 seeded probe indices, within 1e-4 (|fd| + 1); at N = 2^26 the same
   probes and the loss against scihip_reduce of the same chunk;
 - the product adjoint: d left = G right^T and d right = left^T G against float64 numpy, relative
   Frobenius error within the mode's bound, and the gradient of sum((A B - C)^2) against torch.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch  # imported at collection: its first import can be slow on a fresh machine

import libsci_amd as sci
import oracle
from helpers import adjoint_operands, c5_numpy
from libsci_amd.chunk import ARITY, Code, Precision, record_postfix

pytestmark = pytest.mark.gpu

DTYPES = {"fp32": (np.float32, Precision.single), "fp64": (np.float64, Precision.wide)}


def single_code(code):
    return [("in", k) for k in range(ARITY[Code(code)])] + [("op", Code(code))]


def oracle_adjoint(program, arrays, precision, seed=None):
    chunk = record_postfix(program, precision, "reduce", addresses=[a.ctypes.data for a in arrays])
    return oracle.adjoint(chunk, arrays[0].size, seed=seed, dtype=arrays[0].dtype)


def close(got, want, dtype, what):
    relative, absolute = (1e-5, 1e-6) if dtype == np.float32 else (1e-11, 1e-12)
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    bound = relative * np.abs(want) + absolute * max(1.0, np.abs(want).max())
    worst = np.abs(got - want) - bound
    assert (worst <= 0).all(), f"{what}: worst excess {worst.max():.3e} at {int(np.argmax(worst))}"


def c5_expression(a, b, c, d):
    return sci.sinh(a * b + c) * sci.exponential(-(d * d))


C5 = [("in", 0), ("in", 1), ("op", Code.multiply), ("in", 2), ("op", Code.add), ("op", Code.sinh),
      ("in", 3), ("in", 3), ("op", Code.multiply), ("op", Code.negate), ("op", Code.exponential),
      ("op", Code.multiply)]


@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
@pytest.mark.parametrize("kind", ["fp32", "fp64"])
def test_every_code(device, kind, policy):
    dtype, precision = DTYPES[kind]
    device.set_policy(policy)
    try:
        for code in Code:
            print(f"[test] adjoint {code.name} {kind} policy={policy}", flush=True)
            arrays = adjoint_operands(code, 1001, seed=300 + int(code), dtype=dtype)
            operands = [sci.Array.from_numpy(a, device) for a in arrays]
            loss, gradients = sci.gradient(sci.Node(code, operands), operands)
            want_loss, want = oracle_adjoint(single_code(code), arrays, precision)
            terms = oracle.execute(record_postfix(single_code(code), precision, "output",
                                                  addresses=[a.ctypes.data for a in arrays]), arrays[0].size, dtype=dtype)
            tolerance = 1e-6 if dtype == np.float32 else 1e-12
            assert abs(loss - want_loss) <= tolerance * max(1.0, np.abs(terms.astype(np.float64)).sum()), code.name
            for k, g in enumerate(gradients):
                close(g.host(), want[k], dtype, f"{code.name} operand {k}")
    finally:
        device.set_policy(0)


@pytest.mark.parametrize("kind", ["fp32", "fp64"])
def test_seeded_vector_jacobian_products(device, kind):
    dtype, precision = DTYPES[kind]
    n = 777
    seed = np.linspace(-3.0, 2.0, n).astype(dtype)
    upstream = sci.Array.from_numpy(seed, device)
    for code in (Code.tanh, Code.divide, Code.fuse, Code.minimum, Code.absolute):
        arrays = adjoint_operands(code, n, seed=int(code), dtype=dtype)
        operands = [sci.Array.from_numpy(a, device) for a in arrays]
        _, gradients = sci.gradient(sci.Node(code, operands), operands, seed=upstream, loss=False)
        _, want = oracle_adjoint(single_code(code), arrays, precision, seed=seed)
        for k, g in enumerate(gradients):
            close(g.host(), want[k], dtype, f"{code.name} operand {k}")


@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
def test_c5_against_oracle_torch_and_differences(device, policy):
    n = 100_003  # not a multiple of the vector width or the workgroup
    arrays = [oracle.uniform(n, 40 + k, -2.0, 2.0) for k in range(4)]
    operands = [sci.Array.from_numpy(a, device) for a in arrays]
    device.set_policy(policy)
    try:
        loss, gradients = sci.gradient(c5_expression(*operands), operands)
    finally:
        device.set_policy(0)
    got = [g.host() for g in gradients]
    want_loss, want = oracle_adjoint(C5, arrays, Precision.single)
    assert abs(loss - want_loss) <= 1e-6 * np.abs(c5_numpy(*arrays)[0]).sum()
    for k in range(4):
        close(got[k], want[k], np.float32, f"oracle, input {k}")

    # torch's float64 autograd of the same expression at the same (float32) inputs
    leaves = [torch.tensor(a.astype(np.float64), requires_grad=True) for a in arrays]
    reference = (torch.sinh(leaves[0] * leaves[1] + leaves[2]) * torch.exp(-(leaves[3] * leaves[3]))).sum()
    reference.backward()
    assert abs(loss - reference.item()) <= 1e-6 * np.abs(c5_numpy(*arrays)[0]).sum()
    for k in range(4):
        close(got[k], leaves[k].grad.numpy(), np.float32, f"torch, input {k}")

    # float64 central differences at 64 seeded probe indices, step 1e-3
    probes = np.random.default_rng(64).choice(n, 64, replace=False)
    x = [a.astype(np.float64) for a in arrays]
    for k in range(4):
        for i in probes:
            up = [v[i] for v in x]
            down = [v[i] for v in x]
            up[k] += 1e-3
            down[k] -= 1e-3
            difference = (c5_numpy(*up)[0] - c5_numpy(*down)[0]) / 2e-3
            assert abs(got[k][i] - difference) <= 1e-4 * (abs(difference) + 1.0), (k, i)


class ChunkLoss:
    """A torch.autograd.Function over the library: forward runs the fused adjoint kernel (loss and
    gradients in one pass) and backward scales the saved gradients by the upstream gradient."""

    @staticmethod
    def function(device):
        class Function(torch.autograd.Function):
            @staticmethod
            def forward(ctx, *tensors):
                operands = [sci.Array.from_numpy(t.detach().cpu().numpy(), device) for t in tensors]
                loss, gradients = sci.gradient(c5_expression(*operands), operands)
                ctx.save_for_backward(*[torch.from_numpy(g.host()) for g in gradients])
                return torch.tensor(loss, dtype=torch.float64)

            @staticmethod
            def backward(ctx, upstream):
                return tuple(upstream.to(torch.float32) * g for g in ctx.saved_tensors)

        return Function


def test_autograd_function_wrapper(device):
    n = 4096
    values = [oracle.uniform(n, 90 + k, -2.0, 2.0) for k in range(4)]
    leaves = [torch.tensor(v, requires_grad=True) for v in values]
    loss = ChunkLoss.function(device).apply(*leaves)
    (3.0 * loss).backward()
    wide = [torch.tensor(v.astype(np.float64), requires_grad=True) for v in values]
    reference = 3.0 * (torch.sinh(wide[0] * wide[1] + wide[2]) * torch.exp(-(wide[3] * wide[3]))).sum()
    reference.backward()
    assert abs(loss.item() * 3.0 - reference.item()) <= 1e-6 * abs(reference.item()) + 1e-3
    for k in range(4):
        close(leaves[k].grad.numpy(), wide[k].grad.numpy(), np.float32, f"input {k}")


def test_c5_full_size(device):
    """N = 2^26. The loss is checked against the oracle through independent links: a*b + c and
    exp(-(d*d)) equal the oracle's bit for bit, sinh (glibc on the CPU, the device library on the GPU)
    is within 4 ulp of glibc, and the fused kernel's loss is within n * 2^-53 * sum |x| of the exact
    sum of the exact double products of those float factors. The gradients are checked at 64 probes
    against the float64 closed form."""
    from helpers import exact_sum, ulps32

    n = 1 << 26
    print(f"[test] adjoint C5 N={n}", flush=True)
    operands = [sci.uniform(n, device, 40 + k, -2.0, 2.0) for k in range(4)]
    a, b, c, d = operands
    expression = c5_expression(*operands)
    loss, gradients = sci.gradient(expression, operands)
    reduced = sci.reduce_raw(c5_expression(*operands), sci.Fold.sum)
    inner = sci.evaluate(a * b + c).host()
    left = sci.evaluate(sci.sinh(a * b + c)).host()
    right = sci.evaluate(sci.exponential(-(d * d))).host()
    hosts = [oracle.uniform(n, 40 + k, -2.0, 2.0) for k in range(4)]
    assert (inner.view(np.uint32) == oracle.map_f32(Code.fuse, hosts[0], hosts[1], hosts[2],
                                                     threads=16).view(np.uint32)).all()
    square = oracle.map_f32(Code.multiply, hosts[3], hosts[3], threads=16)
    want_right = oracle.map_f32(Code.exponential, oracle.map_f32(Code.negate, square), threads=16)
    assert (right.view(np.uint32) == want_right.view(np.uint32)).all()
    sinh_ulps = int(ulps32(left, oracle.map_f32(Code.sinh, inner, threads=16)).max())
    print(f"[test] sinh over 2^26 inputs: max {sinh_ulps} ulp between the GPU and glibc", flush=True)
    assert sinh_ulps <= 4
    products = left.astype(np.float64) * right.astype(np.float64)
    exact = exact_sum(products)
    bound = n * 2 ** -53 * float(np.abs(products).sum())
    print(f"[test] loss {loss!r} reduced {reduced!r} exact {exact!r} bound {bound:.3g}", flush=True)
    assert abs(loss - exact) <= bound and abs(reduced - exact) <= bound
    probes = np.random.default_rng(26).choice(n, 64, replace=False)
    for i in probes:
        x = [float(o.view(int(i), 1).host()[0]) for o in operands]
        _, want = c5_numpy(*x)
        for k in range(4):
            got = float(gradients[k].view(int(i), 1).host()[0])
            assert abs(got - want[k]) <= 1e-5 * abs(want[k]) + 1e-6 * max(1.0, abs(want[k])), (i, k, got, want[k])


PRODUCT_BOUNDS = {"fp32": 1e-5, "bf16x3": 1e-5, "bf16": 1e-2, "fp64": 1e-13}


@pytest.mark.parametrize("mode", list(PRODUCT_BOUNDS))
@pytest.mark.parametrize("shape", [(64, 64, 64), (100, 51, 37), (257, 300, 129)], ids=str)
def test_product_adjoint(device, mode, shape):
    rows, columns, depth = shape
    dtype = np.float64 if mode == "fp64" else np.float32
    rng = np.random.default_rng(rows * columns + depth)
    a = rng.uniform(-1, 1, (rows, depth)).astype(dtype)
    b = rng.uniform(-1, 1, (depth, columns)).astype(dtype)
    g = rng.uniform(-1, 1, (rows, columns)).astype(dtype)
    da, db = sci.product_adjoint(sci.Array.from_numpy(a, device), sci.Array.from_numpy(b, device),
                                 sci.Array.from_numpy(g, device), mode)
    want_a = g.astype(np.float64) @ b.astype(np.float64).T
    want_b = a.astype(np.float64).T @ g.astype(np.float64)
    for got, want, name in ((da.host(), want_a, "left"), (db.host(), want_b, "right")):
        error = np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want)
        assert error <= PRODUCT_BOUNDS[mode], (name, error)


def test_squared_error_gradient_against_torch(device):
    rows, columns, depth = 192, 160, 320
    rng = np.random.default_rng(5)
    a, b, c = (rng.uniform(-1, 1, s).astype(np.float32) for s in ((rows, depth), (depth, columns), (rows, columns)))
    A, B, C = (sci.Array.from_numpy(x, device) for x in (a, b, c))
    P = sci.product(A, B)
    residual = P - C
    loss, (seed,) = sci.gradient(residual * residual, [P])
    d_a, d_b = sci.product_adjoint(A, B, seed)
    ta, tb = (torch.tensor(x.astype(np.float64), requires_grad=True) for x in (a, b))
    reference = ((ta @ tb - torch.tensor(c.astype(np.float64))) ** 2).sum()
    reference.backward()
    assert abs(loss - reference.item()) <= 1e-5 * reference.item()
    for got, want in ((d_a.host(), ta.grad.numpy()), (d_b.host(), tb.grad.numpy())):
        error = np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want)
        assert error <= 1e-5, error


def test_empty_and_degenerate(device):
    empty = sci.Array(device, (0,), np.float32)
    assert sci.gradient(sci.sin(empty), [empty])[0] == 0.0
    a = sci.Array.from_numpy(np.ones((3, 0), np.float32), device)
    b = sci.Array.from_numpy(np.ones((0, 4), np.float32), device)
    g = sci.Array.from_numpy(np.ones((3, 4), np.float32), device)
    d_a, d_b = sci.product_adjoint(a, b, g)
    assert d_a.shape == (3, 0) and d_b.shape == (0, 4)
