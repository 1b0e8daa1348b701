"""GPU parity at the benchmark workloads' real geometry (tests/workloads.py):
the HIP write path against the reference's own CPU path (oracle/_ref, built
from its sources) on the same input, in both pipelines.

  volume_lz4        the whole 256^3 volume, 64^3 chunks, 4^3-chunk shards, lz4
  multiscale_zstd   the whole 256^3 volume (xor+noise and thirds), 3 levels
  multiscale6_zstd  both volumes as two timepoints, 2x2x2 windows, 6 levels
  sensor_zstd       a 128-frame sample at the full 4096x2304 frame, 64x256x256
                    chunks and 1x8x9-chunk shards
  sweep             a two-epoch sample at the full 1000x1000 frame per dtype x
                    chunk shape (32^3, 64^3, 16x128x128); f16, which the
                    reference's CPU path rejects, against the numpy restatement
  serving           every request of the seeded 24-request mix (a frame sample
                    of each), all submitted to the serving loop at once, each
                    compared with the reference run alone on its input

Equal output means what the reference's cross-validation means by it
(tests/test_cross_validate.c:168-622): the same shard keys and append
metadata; with no codec byte-identical shards; with lz4 / zstd the same filled
index slots, valid CRCs and every chunk decoding (stock liblz4 / libzstd) to
the same bytes.
"""
import numpy as np
import pytest

import chucky_amd as ca
from chucky_amd.api import Server
from oracle import chucky_oracle as O
from oracle import ref
from tests import workloads as W
from tests.helpers import assert_same_shards, decode, geometry, parse_shard, run_gpu, run_ref

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]

PIPELINES = {"plain": ca.PIPELINE_PLAIN, "batched": ca.PIPELINE_BATCHED}


def _check(cfg, data, pipeline, pieces=1, device_input=False, want=None, what="", **options):
    g, _, chunk_bytes = geometry(cfg)
    print(f"[{what}] pipeline={pipeline} device_input={device_input} pieces={pieces} {options}", flush=True)
    gs, gmeta, gcur = run_gpu(cfg, data, pieces=pieces, device_input=device_input,
                              pipeline=PIPELINES[pipeline], **options)
    if want is None:
        want = run_ref(cfg, data, pieces=1)
    rs, rmeta, rcur = want
    assert gcur == rcur
    for lv in gmeta:
        assert gmeta[lv] == rmeta[lv], (lv, gmeta[lv], rmeta[lv])
    assert_same_shards(gs, rs, lambda lv: g[lv][0], cfg.codec, chunk_bytes, what)
    return gs


@pytest.fixture(scope="module")
def volumes():
    return W.multiscale_volumes()


@pytest.mark.parametrize("pipeline", sorted(PIPELINES))
@pytest.mark.parametrize("device_input", [False, True])
def test_volume_lz4_256(volumes, pipeline, device_input):
    cfg = W.volume_lz4()
    _check(cfg, volumes[0], pipeline, pieces=3, device_input=device_input, what="volume_lz4")


@pytest.mark.parametrize("pipeline", sorted(PIPELINES))
@pytest.mark.parametrize("vol", ["xor", "thirds"])
def test_multiscale_zstd_256(volumes, pipeline, vol):
    cfg = W.multiscale_zstd()
    g, _, _ = geometry(cfg)
    assert len(g) == 3
    data = volumes[0 if vol == "xor" else 1]
    _check(cfg, data, pipeline, pieces=2, device_input=(pipeline == "batched"),
           what=f"multiscale_zstd/{vol}")


@pytest.mark.parametrize("pipeline", sorted(PIPELINES))
@pytest.mark.parametrize("pieces", [1, 3])
def test_multiscale_zstd_256_host_untimed(volumes, pipeline, pieces):
    # the benchmark's form: host memory appends (the first piece of a call
    # staged and copied in sub-ranges, the next ones copied ahead), no event
    # markers between the scatter / LOD launches; the batched pipeline
    # compresses the coarse levels of the epochs that emit them in four 2 KiB
    # blocks per 8 KiB segment
    cfg = W.multiscale_zstd()
    _check(cfg, volumes[0], pipeline, pieces=pieces, device_input=False, pass_timing=2,
           what="multiscale_zstd/host")


@pytest.mark.parametrize("block", [4096, 16384])
def test_multiscale_zstd_256_coarse_parts(volumes, block):
    # the coarse levels of split epochs at other segment sizes: 4 KiB segments
    # in two 2 KiB blocks, 16 KiB segments in four 4 KiB blocks
    cfg = W.multiscale_zstd()
    _check(cfg, volumes[1], "batched", pieces=1, device_input=False, pass_timing=2,
           zstd_block_bytes=block, what=f"multiscale_zstd/parts{block}")


@pytest.mark.parametrize("pipeline", sorted(PIPELINES))
def test_multiscale6_zstd_256(volumes, pipeline):
    cfg = W.multiscale6_zstd(timepoints=2)
    g, E, _ = geometry(cfg)
    assert len(g) == 6 and E == 256 ** 3
    _check(cfg, volumes, pipeline, pieces=3, what="multiscale6_zstd")


SENSOR_FRAMES = 128


@pytest.fixture(scope="module")
def sensor_sample():
    cfg = W.sensor_zstd(frames=SENSOR_FRAMES)
    dev = W.sensor_frames_device(SENSOR_FRAMES)
    host = dev.to_host().view(np.uint16)
    dev.free()
    return cfg, host, run_ref(cfg, host, pieces=1)


@pytest.mark.parametrize("pipeline", sorted(PIPELINES))
def test_sensor_zstd_sample(sensor_sample, pipeline):
    cfg, host, want = sensor_sample
    g, E, _ = geometry(cfg)
    assert E == 64 * 4096 * 2304 and g[0][0] == 72  # 1x8x9 chunks per shard
    assert 0 < host.max() <= 4095
    _check(cfg, host, pipeline, pieces=5, want=want, what="sensor_zstd")


def _ulp16(a, b):
    ia = a.view(np.int16).astype(np.int32)
    ib = b.view(np.int16).astype(np.int32)
    ia = np.where(ia < 0, -32768 - ia, ia)
    ib = np.where(ib < 0, -32768 - ib, ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize("case", W.sweep_cases(), ids=lambda c: "-".join(c))
def test_sweep_sample(case):
    dtype, chunk, method, codec = case
    ct = W.SWEEP_CHUNKS[chunk][0]
    frames = 2 * ct
    cfg = W.sweep(dtype, chunk, method, codec, frames=frames)
    g, E, chunk_bytes = geometry(cfg)
    assert len(g) == (4 if chunk == "c16x128" else 5)
    n = frames * 1000 * 1000
    data = W.sweep_data(dtype, n, seed=500 + W.sweep_cases().index(case))
    what = f"sweep/{dtype}/{chunk}/{method}/{codec}"
    if dtype != "f16":
        _check(cfg, data, "plain" if W.sweep_cases().index(case) % 2 else "batched", pieces=2, what=what)
        return
    # f16: the restatement follows the reference's GPU arithmetic (float
    # accumulator for the mean); the GPU must equal it bit for bit
    print(f"[{what}] f16 against the restatement", flush=True)
    gs, _, gcur = run_gpu(cfg, data, pieces=2, pipeline=ca.PIPELINE_BATCHED)
    dims = [O.ODim(d.size, d.chunk, d.cps, d.downsample) for d in cfg.dims]
    want, wcur = O.write_stream(dims, np.float16, data, reduce_method=ca.REDUCE_ID[method], max_nlod=5)
    assert gcur == wcur
    assert set(gs) == set(want)
    worst = 0
    for k in want:
        gi, wi = parse_shard(gs[k], g[k[0]][0]), parse_shard(want[k], g[k[0]][0])
        assert [e is None for e in gi] == [e is None for e in wi], k
        for ge, we in zip(gi, wi):
            if we is None:
                continue
            a = np.frombuffer(decode(codec, gs[k][ge[0]:ge[0] + ge[1]], chunk_bytes), dtype=np.float16)
            b = np.frombuffer(want[k][we[0]:we[0] + we[1]], dtype=np.float16)
            assert a.shape == b.shape
            worst = max(worst, int(_ulp16(a, b).max()))
    # the stated tolerance is 1 ULP; the GPU follows the same operations in
    # the same order, so it is expected (and required) to be exact
    assert worst == 0, f"{what}: f16 differs by {worst} ULP"


SERVING_EXTRA_FRAMES = 4


def _serving_sample(req):
    _, dt, shape, chunk, codec, nlod = req
    frames = chunk[0] + SERVING_EXTRA_FRAMES  # one epoch and a partial one
    return W.serving_config(dt, (frames,) + tuple(shape[1:]), chunk, codec, nlod), (frames,) + tuple(shape[1:])


def test_serving_mix_matches_reference():
    reqs = W.serving_requests()
    inputs = {}
    for dt, shape in W.SERVING_SHAPES:
        frames = max(r[3][0] for r in reqs if r[1] == dt) + SERVING_EXTRA_FRAMES
        n = frames * int(np.prod(shape[1:]))
        inputs[dt] = W.sweep_data(dt, n, seed=900 + ca.DTYPE_ID[dt])
    # the reference, once per distinct request type
    want = {}
    cfgs = []
    for r in reqs:
        cfg, shape = _serving_sample(r)
        cfgs.append((cfg, shape))
        key = r[1:]
        if key not in want:
            data = inputs[r[1]][:int(np.prod(shape))]
            print(f"[serving] reference for {key}", flush=True)
            want[key] = run_ref(cfg, data, pieces=1)
    assert len(want) >= 6
    srv = Server(hbm_budget_bytes=200 * 10**9, max_concurrent=24, pipeline=ca.PIPELINE_BATCHED)
    sinks = []
    for r, (cfg, shape) in zip(reqs, cfgs):
        sinks.append(ca.MemSink())
        srv.submit(cfg, sinks[-1], inputs[r[1]][:int(np.prod(shape))])
    assert srv.wait() == 0, ca.last_error()
    reps = srv.reports()
    running, _ = srv.peaks()
    srv.close()
    assert len(reps) == len(reqs) and all(x.status == 0 for x in reps)
    assert running >= 2, "requests never overlapped"
    for i, (r, (cfg, _)) in enumerate(zip(reqs, cfgs)):
        g, _, chunk_bytes = geometry(cfg)
        rs, rmeta, _ = want[r[1:]]
        got = sinks[i].shards()
        for lv in g:
            assert sinks[i].append_sizes(lv) == rmeta[lv], (i, lv)
        assert_same_shards(got, rs, lambda lv: g[lv][0], cfg.codec, chunk_bytes, f"serving request {i} {r[1:]}")
