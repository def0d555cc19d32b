"""On-wire request/result frames and JSON records (host only)."""
import json
import struct

import numpy as np
import pytest

from blockstructuredsolvers_amd import wire, workloads
from blockstructuredsolvers_amd.serving import Request, Result


def dense_request(N=5, n=3, nrhs=2, k=2, seed=0, rid=7):
    rng = np.random.default_rng(seed)
    A = np.asfortranarray(rng.standard_normal((n, n, N)))
    B = np.asfortranarray(rng.standard_normal((n, n, N - 1)))
    rhs = [np.asfortranarray(rng.standard_normal((n, nrhs, N))) for _ in range(k)]
    return Request(rid=rid, arrival=0.25, rhs=rhs, A=A, B=B)


def test_dense_round_trip_f64():
    q = dense_request()
    out = wire.decode_request(wire.encode_request(q))
    assert (out.rid, out.arrival, out.N, out.n, out.nrhs) == (q.rid, q.arrival, q.N, q.n, q.nrhs)
    assert out.model is None
    assert np.array_equal(out.A, q.A) and np.array_equal(out.B, q.B)
    assert len(out.rhs) == 2 and all(np.array_equal(a, b) for a, b in zip(out.rhs, q.rhs))
    assert out.A.flags.f_contiguous and out.A.flags.writeable


def test_dense_round_trip_f32_and_single_block():
    q = dense_request(N=1, n=4, nrhs=1, k=1)
    q.B = None  # one block has no coupling
    buf = wire.encode_request(q, dtype=np.float32)
    assert len(buf) == 48 + 4 * (4 * 4 * 1 + 4 * 1 * 1)
    out = wire.decode_request(buf)
    assert out.A.dtype == np.float32 and out.B.shape == (4, 4, 0)
    assert np.array_equal(out.A, q.A.astype(np.float32))


def test_model_round_trip():
    rng = np.random.default_rng(3)
    model = workloads.smoother_model(16, rng)
    d = np.asfortranarray(rng.standard_normal((16, 1, 40)))
    q = Request(rid=2 ** 40 + 5, arrival=1.5, rhs=[d], model=model)
    buf = wire.encode_request(q)
    m = model.H.shape[0]
    assert len(buf) == 48 + 8 * (2 * 16 * 16 + m * 16 + m * m) + 8 * 16 * 40
    out = wire.decode_request(buf)
    assert out.rid == q.rid and out.A is None and out.B is None
    for name in ("F", "H", "Qinv", "Rinv"):
        assert np.array_equal(getattr(out.model, name), getattr(model, name))
    A0, B0 = out.dense()
    assert np.allclose(A0[:, :, 7], model.blocks()[0])


def test_request_header_layout():
    """Byte layout of the header, written out by hand: little-endian, 48 bytes."""
    q = Request(rid=0x0102030405060708, arrival=2.0,
                rhs=[np.zeros((2, 1, 3), order="F")], A=np.zeros((2, 2, 3), order="F"),
                B=np.zeros((2, 2, 2), order="F"))
    head = wire.encode_request(q)[:48]
    expect = (b"BSSQ" + b"\x01\x00" + b"\x00\x00"
              + bytes([8, 7, 6, 5, 4, 3, 2, 1])
              + bytes([0, 0, 0, 0, 0, 0, 0, 0x40])          # 2.0 as f64
              + b"\x03\x00\x00\x00" + b"\x02\x00\x00\x00"   # N, n
              + b"\x01\x00\x00\x00" + b"\x01\x00\x00\x00"   # nrhs, k
              + b"\x00\x00\x00\x00" + b"\x00" + b"\x00\x00\x00")  # m, dtype, reserved
    assert head == expect


def test_payload_is_column_major_blocks():
    """Element (r, c) of block i sits at (i*n*n + c*n + r), as in create_matrix_list."""
    n, N = 3, 2
    A = np.zeros((n, n, N), order="F")
    A[1, 2, 1] = 5.0  # row 1, column 2, block 1
    q = Request(rid=1, arrival=0.0, rhs=[np.zeros((n, 1, N), order="F")], A=A,
                B=np.zeros((n, n, N - 1), order="F"))
    buf = wire.encode_request(q)
    vals = np.frombuffer(buf, dtype="<f8", count=n * n * N, offset=48)
    assert np.flatnonzero(vals).tolist() == [1 * n * n + 2 * n + 1]


@pytest.mark.parametrize("mutate, msg", [
    (lambda b: b"XXXX" + b[4:], "magic"),
    (lambda b: b[:4] + struct.pack("<H", 9) + b[6:], "version"),
    (lambda b: b[:6] + struct.pack("<H", 6) + b[8:], "flags"),
    (lambda b: b[:44] + b"\x07" + b[45:], "dtype"),
    (lambda b: b[:24] + struct.pack("<I", 0) + b[28:], ">= 1"),
    (lambda b: b[:-8], "truncated"),
    (lambda b: b + b"\x00" * 8, "trailing"),
    (lambda b: b[:20], "shorter"),
    (lambda b: b[:40] + struct.pack("<I", 3) + b[44:], "observation size"),
])
def test_malformed_request_frames(mutate, msg):
    buf = wire.encode_request(dense_request())
    with pytest.raises(ValueError, match=msg):
        wire.decode_request(mutate(buf))


def test_result_round_trip_and_record():
    rng = np.random.default_rng(5)
    xs = [np.asfortranarray(rng.standard_normal((4, 2, 6))) for _ in range(3)]
    res = Result(rid=11, x=xs, arrival=0.5, admitted=0.75, finished=1.0,
                 solve_ms=[0.1, 0.2, 0.3], info=0)
    rid, info, out = wire.decode_result(wire.encode_result(res))
    assert (rid, info) == (11, 0) and all(np.array_equal(a, b) for a, b in zip(out, xs))
    rec = json.loads(json.dumps(wire.result_record(res)))
    assert rec["schema"] == wire.RESULT_SCHEMA
    assert (rec["N"], rec["n"], rec["nrhs"], rec["rhs"]) == (6, 4, 2, 3)
    assert rec["latency_ms"] == pytest.approx(500.0)
    bad = wire.encode_result(res)
    with pytest.raises(ValueError, match="magic"):
        wire.decode_result(b"BSSQ" + bad[4:])
    with pytest.raises(ValueError, match="truncated"):
        wire.decode_result(bad[:-1])


def test_benchmark_record_matches_reference_fields():
    """init/factorize/solve/total in ms as mean and sample deviation (Julia's std),
    residual likewise (benchmark/benchmark_gpu.jl:227-239)."""
    timings = [(0.001, 0.002, 0.003), (0.003, 0.004, 0.005)]
    rec = wire.benchmark_record(100, 32, "BlockTriDiag", timings, [1e-14, 3e-14])
    assert rec["total_size"] == 3200 and rec["iterations"] == 2
    assert rec["init_ms"]["mean"] == pytest.approx(2.0)
    assert rec["factorize_ms"]["mean"] == pytest.approx(3.0)
    assert rec["total_ms"]["mean"] == pytest.approx(7.0)       # (2+3 + 4+5) / 2
    assert rec["solve_ms"]["std"] == pytest.approx(np.sqrt(2.0))  # sample std of {3, 5}
    assert rec["residual"]["mean"] == pytest.approx(2e-14)
    json.dumps(rec)
