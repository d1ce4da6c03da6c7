"""GPU tests of the render-job server (cwp_server_*): batched multi-job passes and continuous batching.

A job's image is a pure function of the job (scene, resolution, samples, bounce limit, seed) and of its samples per
pass: each path's arithmetic and each pixel's summation order are those of the job rendered alone. So images are
compared bit-exact between a job alone, the same job in a mix, and the plain one-job-at-a-time mode; a job at the
scene's own resolution is bit-identical to the single-job renderer's image at the same samples per pass, and reaches
>= 50 dB PSNR against the oracle.
"""
import math

import numpy as np
import pytest

import cwp
import pyoracle as po

pytestmark = pytest.mark.gpu

TERRAIN = ["--gen", "terrain", "--gen_grid", "40", "--gen_boxes", "20", "--width", "64", "--height", "48"]
BOXES = ["--gen", "boxes", "--gen_boxes", "64", "--width", "32", "--height", "32", "--seed", "7"]

# (scene, width, height, spp, bounces, seed): ragged sizes, odd sample counts, different bounce limits
JOBS = [(0, 64, 48, 4, 4, 1), (1, 32, 32, 8, 3, 2), (0, 50, 30, 6, 4, 3), (1, 64, 64, 2, 2, 4), (0, 17, 9, 5, 4, 5),
        (1, 1, 1, 3, 4, 6), (0, 64, 48, 4, 0, 7)]


def psnr(a, b):
    def srgb(x):
        x = np.clip(x, 0, 1)
        return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)
    mse = float(np.mean((srgb(a.astype(np.float64)) - srgb(b.astype(np.float64))) ** 2))
    return 99.0 if mse == 0 else 10 * math.log10(1.0 / mse)


def setups(variant=8):
    return [cwp.Setup(a + ["--variant", "bvh%d" % variant, "--bounces", "4"]) for a in (TERRAIN, BOXES)]


def run_jobs(srv, jobs, batched, arrivals=None):
    srv.clear_jobs()
    ids = [srv.submit(*j[:5], seed=j[5], arrival_s=(arrivals[k] if arrivals else 0.0)) for k, j in enumerate(jobs)]
    st = srv.run(batched=batched)
    return [srv.image(i) for i in ids], [srv.job(i) for i in ids], st


@pytest.mark.parametrize("variant,max_paths,graph,rpl,rolling", [(8, 1 << 20, 1, 1, 1), (8, 1 << 14, 1, 1, 1),
                                                                  (8, 1 << 14, 0, 4, 1), (4, 1 << 20, 1, 2, 1),
                                                                  (2, 1 << 14, 1, 1, 1), (8, 1 << 14, 1, 1, 0),
                                                                  (8, 1 << 20, 0, 1, 0)])
def test_job_alone_equals_job_in_mix(variant, max_paths, graph, rpl, rolling):
    """Each job alone, all jobs in one batched run, and the plain mode give bit-identical images; with 2^14 path
    slots the larger jobs take several rounds of samples and wait for slots to free up. Rolling steps (paths of every
    depth and new camera rays in one launch) and whole passes, launches replayed from a captured graph or issued one
    kernel at a time, and any rays per lane, trace the same paths."""
    print("case", variant, max_paths, graph, rpl, rolling, flush=True)
    S = setups(variant)
    srv = cwp.Server(variant=variant, bounces=4, max_paths=max_paths, use_graph=graph, rays_per_lane=rpl, rolling=rolling)
    for s in S:
        srv.add_scene(s)
    alone = []
    for j in JOBS:
        img, info, _ = run_jobs(srv, [j], True)
        alone.append(img[0])
        assert info[0]["passes"] * info[0]["samples_per_pass"] == j[3]
    mix, infos, st = run_jobs(srv, JOBS, True)
    assert st["jobs"] == len(JOBS)
    for k, (a, b) in enumerate(zip(alone, mix)):
        np.testing.assert_array_equal(a, b, err_msg="job %d" % k)
    plain, _, stp = run_jobs(srv, JOBS, False)
    for k, (a, b) in enumerate(zip(alone, plain)):
        np.testing.assert_array_equal(a, b, err_msg="plain job %d" % k)
    assert st["rays"] == stp["rays"]             # the same paths, traced in a different grouping
    assert stp["passes"] >= st["passes"]
    if max_paths == 1 << 14:
        assert max(i["passes"] for i in infos) > 1
        assert st["max_paths_in_flight"] <= max_paths


def test_job_matches_single_job_renderer_and_oracle():
    """A job at the scene's own resolution, bounce limit and seed equals the single-job renderer's image at the same
    samples per pass bit for bit, and matches the oracle's rendering at >= 50 dB."""
    spp, bounces = 8, 4
    base = TERRAIN + ["--bounces", str(bounces), "--spp", str(spp), "--batch", str(spp), "--seed", "1000"]
    s = cwp.Setup(base)
    r = cwp.Renderer(s)
    r.render(spp)
    ref = r.accum()
    r.close()
    srv = cwp.Server(variant=8, bounces=bounces, max_paths=1 << 20)
    sid = srv.add_scene(s)
    img, info, st = run_jobs(srv, [(sid, 64, 48, spp, bounces, 1000)], True)
    assert info[0]["samples_per_pass"] == spp
    np.testing.assert_array_equal(img[0], ref)
    sc, _ = po.scene_from_setup(s)
    acc, _ = sc.render(s.params(), spp, batch=spp, bounces=bounces)
    assert psnr(img[0][..., :3] / spp, acc[..., :3] / spp) >= 50.0


def test_lru_eviction_keeps_results():
    """A scene budget that holds one scene at a time: jobs alternate scenes, scenes are evicted and re-uploaded, and
    every image is unchanged."""
    S = setups(8)
    probe = cwp.Server(variant=8, bounces=4, max_paths=1 << 10)
    for s in S:
        probe.add_scene(s)
    sizes = list(probe.scene_bytes.values())
    probe.close()
    assert max(sizes) < sum(sizes)
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 20, scene_budget_mb=max(sizes) * 1.01 / 1048576.0)
    for s in S:
        srv.add_scene(s)
    alone = [run_jobs(srv, [j], True)[0][0] for j in JOBS[:4]]
    mix, _, st = run_jobs(srv, JOBS[:4], False)
    assert st["scenes_evicted"] >= 2
    for a, b in zip(alone, mix):
        np.testing.assert_array_equal(a, b)


def test_arrivals_and_latency():
    """Jobs arriving over time are admitted when they arrive; latency = finish - arrival; percentiles ordered."""
    S = setups(8)
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 20)
    for s in S:
        srv.add_scene(s)
    arr = [0.0, 0.01, 0.02, 0.05, 0.05, 0.08, 0.1]
    imgs, infos, st = run_jobs(srv, JOBS, True, arrivals=arr)
    alone = [run_jobs(srv, [j], True)[0][0] for j in JOBS]
    for a, b in zip(alone, imgs):
        np.testing.assert_array_equal(a, b)
    for i, a in zip(infos, arr):
        assert i["admitted_s"] >= a and i["finished_s"] >= i["admitted_s"]
        assert i["latency_s"] == pytest.approx(i["finished_s"] - a)
    assert 0 < st["latency_p50_s"] <= st["latency_p99_s"]
    assert st["wall_s"] >= 0.1 and st["rays_per_s"] > 0


def test_server_errors():
    S = setups(8)
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 12)
    srv.add_scene(S[0])
    with pytest.raises(cwp.CwpError):
        srv.submit(5, 8, 8, 1)                 # unknown scene
    with pytest.raises(cwp.CwpError):
        srv.submit(0, 8, 8, 1, bounces=9)      # deeper than the server's bounce limit
    with pytest.raises(cwp.CwpError):
        srv.submit(0, 128, 128, 1)             # more pixels than path slots
    jid = srv.submit(0, 8, 8, 1, bounces=2)
    with pytest.raises(cwp.CwpError):
        srv.image(jid)                         # not finished
    with pytest.raises(cwp.CwpError):
        cwp.Server(variant=8, stack_lds=12)
    with pytest.raises(cwp.CwpError):
        cwp.Server(variant=8, rays_per_lane=3)


def test_scene_of_another_layout_is_relaid():
    """A bvh4 setup added to a BVH8 server is rebuilt in the server's layout from its mesh: its jobs equal the jobs of
    the bvh8 setup of the same scene."""
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 20)
    a = srv.add_scene(cwp.Setup(TERRAIN + ["--variant", "bvh8", "--bounces", "4"]))
    b = srv.add_scene(cwp.Setup(TERRAIN + ["--variant", "bvh4", "--bounces", "4"]))
    assert srv.scene_bytes[a] == srv.scene_bytes[b]
    imgs, _, _ = run_jobs(srv, [(a, 64, 48, 4, 4, 1), (b, 64, 48, 4, 4, 1)], True)
    np.testing.assert_array_equal(imgs[0], imgs[1])


def test_autotune_picks_a_configuration_and_keeps_results():
    """The autotuner runs the mix under every (width, stack, rays per lane, compaction granularity), keeps the fastest,
    leaves the jobs unfinished; the next run under the chosen configuration gives the same images (bit-exact for the
    BVH8 layout, >= 50 dB for BVH4, whose closest hits can differ only among triangles at equal distance)."""
    S = setups(8)
    ref_srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 20)
    for s in S:
        ref_srv.add_scene(s)
    ref, _, _ = run_jobs(ref_srv, JOBS, True)
    ref_srv.close()
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 20)
    for s in S:
        srv.add_scene(s)
    srv.clear_jobs()
    ids = [srv.submit(*j[:5], seed=j[5]) for j in JOBS]
    trials, chosen = srv.autotune(variants=(4, 8), stacks=(8, 16), rays_per_lane=(1, 2), compacts=(0, 1))
    assert len(trials) == 16 and all(t["rays_per_s"] > 0 for t in trials)
    assert {t["compact"] for t in trials} == {0, 1}
    best = max(trials, key=lambda t: t["rays_per_s"])
    assert (chosen["variant"], chosen["stack_lds"], chosen["rays_per_lane"], chosen["compact"]) == \
        (best["variant"], best["stack_lds"], best["rays_per_lane"], best["compact"])
    assert all(srv.job(i)["finished_s"] < 0 for i in ids)
    srv.run(batched=True)
    for i, r in zip(ids, ref):
        img = srv.image(i)
        if chosen["variant"] == 8:
            np.testing.assert_array_equal(img, r)
        else:
            spp = srv.job(i)["spp"]
            assert psnr(img[..., :3] / spp, r[..., :3] / spp) >= 50.0


def test_full_integrator_jobs_alone_equal_mix():
    """full-integrator server: materials, emitters and an environment; each job alone equals the job in the mix and in
    plain mode, and a job at the scene's resolution equals the single-job renderer's image"""
    mats = ["--gen_materials", "4", "--gen_env", "64,32,2,5"]
    S = [cwp.Setup(TERRAIN + mats + ["--bounces", "4"]), cwp.Setup(BOXES + ["--gen_materials", "6", "--bounces", "4"])]
    assert all(s.info["full"] for s in S)
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 16, full=1)
    for s in S:
        srv.add_scene(s)
    alone = [run_jobs(srv, [j], True)[0][0] for j in JOBS]
    mix, _, _ = run_jobs(srv, JOBS, True)
    plain, _, _ = run_jobs(srv, JOBS, False)
    for k in range(len(JOBS)):
        np.testing.assert_array_equal(alone[k], mix[k], err_msg="job %d" % k)
        np.testing.assert_array_equal(alone[k], plain[k], err_msg="plain job %d" % k)
    spp = 8
    s = cwp.Setup(TERRAIN + mats + ["--bounces", "4", "--spp", str(spp), "--batch", str(spp), "--seed", "1000"])
    r = cwp.Renderer(s)
    r.render(spp)
    ref = r.accum()
    r.close()
    srv2 = cwp.Server(variant=8, bounces=4, max_paths=1 << 20, full=1)
    sid = srv2.add_scene(s)
    img, info, _ = run_jobs(srv2, [(sid, 64, 48, spp, 4, 1000)], True)
    np.testing.assert_array_equal(img[0], ref)


def test_legacy_scene_on_full_server_and_full_scene_on_legacy_server():
    full = cwp.Setup(TERRAIN + ["--gen_materials", "4", "--bounces", "4"])
    legacy = cwp.Setup(BOXES + ["--bounces", "4"])
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 16)
    with pytest.raises(cwp.CwpError):
        srv.add_scene(full)
    fsrv = cwp.Server(variant=8, bounces=4, max_paths=1 << 16, full=1)
    sid = fsrv.add_scene(legacy)
    img, _, _ = run_jobs(fsrv, [(sid, 32, 32, 8, 3, 2)], True)
    # the full integrator on Lambert materials estimates the same image as the legacy one (different sampling
    # arithmetic, same expectation): compare with the legacy server at 8 spp loosely
    lsrv = cwp.Server(variant=8, bounces=4, max_paths=1 << 16)
    lid = lsrv.add_scene(legacy)
    limg, _, _ = run_jobs(lsrv, [(lid, 32, 32, 8, 3, 2)], True)
    assert psnr(img[0][..., :3] / 8, limg[0][..., :3] / 8) >= 40.0


def test_rolling_admits_arrivals_at_the_next_step():
    """A long job (several rounds of samples) is running when a small job arrives: with rolling steps the small job
    starts at the next launch and is done long before the long one; both images equal their images alone."""
    S = setups(8)
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 18)
    for s in S:
        srv.add_scene(s)
    big, small = (0, 256, 256, 64, 4, 11), (1, 16, 16, 4, 3, 12)
    alone = [run_jobs(srv, [j], True)[0][0] for j in (big, small)]
    imgs, infos, st = run_jobs(srv, [big, small], True, arrivals=[0.0, 0.01])
    np.testing.assert_array_equal(imgs[0], alone[0])
    np.testing.assert_array_equal(imgs[1], alone[1])
    assert infos[0]["passes"] > 1
    assert infos[1]["finished_s"] < infos[0]["finished_s"]
    assert infos[1]["latency_s"] < 0.5 * infos[0]["latency_s"], infos
    assert st["steps"] > 0


def test_per_lane_compaction_gives_the_same_images():
    """Rolling steps with per-lane queue compaction (one atomic per appended ray, so another slot order) render the
    same images bit for bit as per-wave compaction: a path's result depends on its id, never on its slot."""
    S = setups(8)
    out = []
    for cm in (0, 1):
        srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 16, compact=cm)
        assert srv.get_options()["compact"] == cm
        for s in S:
            srv.add_scene(s)
        out.append(run_jobs(srv, JOBS, True)[0])
        srv.close()
    for k in range(len(JOBS)):
        np.testing.assert_array_equal(out[0][k], out[1][k], err_msg="job %d" % k)
    with pytest.raises(cwp.CwpError, match="compaction"):
        cwp.Server(variant=8, bounces=4, max_paths=1 << 16, compact=2)


@pytest.mark.parametrize("full,rolling,compact", [(0, 1, 0), (0, 1, 1), (0, 0, 0), (1, 1, 0), (1, 0, 0)])
def test_straggler_hand_off_gives_the_same_images(full, rolling, compact):
    """The hand-off in batches (a wave with fewer than tail_lanes lanes still tracing writes their traversals as tail
    records; a tail launch resumes them with each record's job and scene) renders the same images bit for bit as no
    hand-off, at every threshold: 64 hands off at the first finished lane, 2 only the last lane; with one level and
    with two (the first tail launch hands its own stragglers on). Rolling steps (camera rays of new jobs resumed by
    camera index) and whole passes, both compactions, both integrators; 2^14 path slots make the larger jobs take
    several rounds."""
    mats = ["--gen_materials", "4", "--gen_env", "64,32,2,5"] if full else []
    S = [cwp.Setup(TERRAIN + mats + ["--bounces", "4"]), cwp.Setup(BOXES + (["--gen_materials", "6"] if full else []) +
                                                               ["--bounces", "4"])]
    out = {}
    for tail, levels in ((0, 2), (2, 2), (12, 2), (64, 2), (12, 1), (64, 1)):
        print("case hand-off", full, rolling, compact, tail, levels, flush=True)
        srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 14, full=full, rolling=rolling, compact=compact, tail_lanes=tail,
                         tail_levels=levels)
        assert (srv.get_options()["tail_lanes"], srv.get_options()["tail_levels"]) == (tail, levels)
        for s in S:
            srv.add_scene(s)
        imgs, infos, st = run_jobs(srv, JOBS, True)
        out[tail, levels] = (imgs, st["rays"])
        if tail == 12:
            plain, _, _ = run_jobs(srv, JOBS, False)
            for k in range(len(JOBS)):
                np.testing.assert_array_equal(plain[k], imgs[k], err_msg="plain job %d" % k)
        srv.close()
    for tail in ((2, 2), (12, 2), (64, 2), (12, 1), (64, 1)):
        assert out[tail][1] == out[0, 2][1]
        for k in range(len(JOBS)):
            np.testing.assert_array_equal(out[tail][0][k], out[0, 2][0][k], err_msg="tail %s job %d" % (tail, k))


def test_live_submissions_from_threads_while_jobs_run():
    """cwp_server_start / cwp_job_submit / cwp_job_poll / cwp_job_result: two threads submit jobs while others are on
    the GPU; every job reaches DONE and its image equals the job rendered alone."""
    import threading
    import time
    S = setups(8)
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 16)
    for s in S:
        srv.add_scene(s)
    alone = [run_jobs(srv, [j], True)[0][0] for j in JOBS]
    srv.clear_jobs()
    srv.start()
    ids = [None] * len(JOBS)

    def submit(ks):
        for k in ks:
            j = JOBS[k]
            ids[k] = srv.submit_live(*j[:5], seed=j[5])
            time.sleep(0.003)
    th = [threading.Thread(target=submit, args=(list(range(0, len(JOBS), 2)),)),
          threading.Thread(target=submit, args=(list(range(1, len(JOBS), 2)),))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    deadline = time.time() + 60
    while any(srv.poll(i) != srv.JOB_DONE for i in ids) and time.time() < deadline:
        time.sleep(0.002)
    states = [srv.poll(i) for i in ids]
    srv.stop()
    assert all(x == srv.JOB_DONE for x in states), states
    for k, j in enumerate(JOBS):
        np.testing.assert_array_equal(srv.result(ids[k], j[1], j[2]), alone[k], err_msg="job %d" % k)
    with pytest.raises(cwp.CwpError):
        srv.submit_live(0, 8, 8, 1)   # stopped


def test_production_scale_scenes_alone_equal_mix_and_renderer():
    """The serving configurations at full scene size: the 118,562-triangle terrain with generated materials, a
    512x256 log-normal environment and generated texture maps (full integrator), beside the 1,024-box scene; jobs of
    ragged resolutions up to 320x180, 5-16 samples, 6 bounces. Each job alone equals the job in the mix and in plain
    mode bit for bit, and the job at the terrain's own resolution equals the single-job renderer's image."""
    full = ["--gen", "terrain", "--width", "320", "--height", "180", "--gen_materials", "3", "--gen_env", "512,256,1.5,2",
            "--gen_textures", "64", "--bounces", "6"]
    boxes = ["--gen", "boxes", "--width", "128", "--height", "128", "--gen_materials", "5", "--bounces", "6"]
    S = [cwp.Setup(full), cwp.Setup(boxes)]
    assert S[0].info["num_tris"] > 100000 and S[0].info["textured"] and all(s.info["full"] for s in S)
    jobs = [(0, 320, 180, 16, 6, 11), (1, 128, 128, 8, 6, 12), (0, 97, 61, 5, 6, 13), (1, 33, 17, 7, 6, 14),
            (0, 320, 180, 16, 6, 15)]
    srv = cwp.Server(variant=8, bounces=6, max_paths=1 << 20, full=1)
    for s in S:
        srv.add_scene(s)
    alone = [run_jobs(srv, [j], True)[0][0] for j in jobs]
    mix, info, _ = run_jobs(srv, jobs, True)
    plain, _, _ = run_jobs(srv, jobs, False)
    for k in range(len(jobs)):
        print("case production job", k, flush=True)
        np.testing.assert_array_equal(alone[k], mix[k], err_msg="job %d" % k)
        np.testing.assert_array_equal(alone[k], plain[k], err_msg="plain job %d" % k)
        assert np.isfinite(mix[k]).all() and mix[k][..., :3].mean() > 0
    spp = info[0]["samples_per_pass"]
    assert 16 % spp == 0
    s = cwp.Setup(full + ["--spp", "16", "--batch", str(spp), "--seed", "11"])
    r = cwp.Renderer(s)
    r.render(16)
    ref = r.accum()
    r.close()
    np.testing.assert_array_equal(mix[0], ref)
