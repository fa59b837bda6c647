"""Serving loop with continuous batching over one serving world.

Requests arrive over time, each a scene of the batched-scene distribution plus
a step count.  A request is admitted as soon as a free body range of the
serving world can hold it (admission by footprint: the world's body capacity
is its HBM budget), its state stays resident on the device, and every tick
advances all admitted scenes by one step in one batched launch per stage.
Finished scenes are read back and their range is freed for the next request.
Scenes never interact and occupy fixed body ranges, so a scene's result is
the same alone and in the mix.
"""
import collections
import math
import time

import numpy as np

from . import scenes

Request = collections.namedtuple("Request", "rid arrival scene bodies steps")


def poisson_requests(n, rate, seed, max_steps=240, lo=8, hi=2048):
    """n requests with exponential inter-arrival times (rate per second); scene
    index rid (per-scene seed base + rid), body count log-uniform [lo, hi],
    step count round(exp(U[0, ln max_steps]))."""
    rng = np.random.default_rng(seed ^ 0xC4)
    gaps = rng.exponential(1.0 / rate, n)
    arrival = np.cumsum(gaps) - gaps[0]
    sizes = scenes.c3_sizes(n, seed, lo, hi)
    steps = np.rint(np.exp(rng.uniform(0.0, math.log(max_steps), n))).astype(int)
    steps = np.maximum(steps, 1)
    return [Request(i, float(arrival[i]), i, int(sizes[i]), int(steps[i])) for i in range(n)], sizes


class SceneSource:
    """Scene payloads for a serving world: the shape libraries are registered
    once; each request's scene (container walls, then its dynamic bodies, as
    scenes.build_batched orders them) references them."""

    def __init__(self, world, sizes, seed, hull_lib_size=256):
        self.w, self.sizes, self.seed = world, sizes, seed
        self.lib = scenes.hull_library(seed, hull_lib_size)
        self.sph_r, self.sph_mp = scenes.sphere_library(seed, hull_lib_size)
        geo, self.cyl_mp, self.cyl_r = scenes.cylinder_library(seed, hull_lib_size)
        self.lib_sid = np.array([world.add_hull_exact(it["body"]) for it in self.lib.items])
        self.sph_sid = np.array([world.add_sphere((0.0, 0.0, 0.0), r) for r in self.sph_r])
        self.cyl_sid = np.array([world.add_cylinder(g) for g in geo])
        self.boxes = {}

    def _box(self, dims):
        if dims not in self.boxes:
            h = scenes.box_body(*dims, 1.0)
            self.boxes[dims] = (self.w.add_hull_exact(h), h["mass_props"])
        return self.boxes[dims]

    def payload(self, s):
        """(shape ids, movement, mass props, material, state, dynamic bodies) of scene s."""
        kind, picks, L, pos, q, vel = scenes.scene_bodies(self.sizes, self.seed, s, self.lib, self.sph_r, self.cyl_r)
        sid, mv, mp, st = [], [], [], []
        for dims, c in scenes.container_walls(L):
            b, m = self._box(dims)
            sid.append(b)
            mv.append(1)
            mp.append(m)
            st.append(np.concatenate([c, [1, 0, 0, 0], [0, 0, 0], [0, 0, 0]]))
        n = len(picks)
        shp = np.select([kind == 0, kind == 1], [self.lib_sid[picks], self.sph_sid[picks]], self.cyl_sid[picks])
        mps = np.select([(kind == 0)[:, None], (kind == 1)[:, None]], [self.lib.mp[picks], self.sph_mp[picks]], self.cyl_mp[picks])
        sid = np.concatenate([np.array(sid, np.int32), shp.astype(np.int32)])
        mv = np.concatenate([np.array(mv, np.int32), np.zeros(n, np.int32)])
        mp = np.concatenate([np.array(mp), mps])
        mat = np.tile(scenes.DEFAULT_MATERIAL, (len(sid), 1))
        st = np.concatenate([np.array(st), np.concatenate([pos, q, vel, np.zeros((n, 3))], axis=1)])
        return sid, mv, mp, mat, st, n


def serve(world, source, requests, keep_states=False, log=None):
    """Runs the requests through the serving world in real time (arrival times
    relative to the start).  Returns per-request latency (arrival to result),
    per-tick step latency, throughput and, optionally, the final states."""
    payloads = {r.rid: source.payload(r.scene) for r in requests}  # the scene data arrives with the request
    cap = world.serving_info()[0]
    too_big = [r.rid for r in requests if len(payloads[r.rid][0]) > cap]
    if too_big:
        raise ValueError(f"requests {too_big[:5]} need more bodies than the serving world holds ({cap})")
    pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
    queue = []
    active = {}
    lat, ticks, waits = {}, [], {}
    states = {}
    diverged = []  # requests whose scene went non-finite (reported, not retried)
    body_steps = 0
    tick_log = []  # (start time, body-steps of the tick)
    t0 = time.perf_counter()
    while pending or queue or active:
        now = time.perf_counter() - t0
        while pending and pending[0].arrival <= now:
            queue.append(pending.popleft())
        if not active and not queue:
            time.sleep(max(0.0, pending[0].arrival - now))
            continue
        still = []
        for r in queue:  # first fit in arrival order; a request that does not fit waits for freed ranges
            sid, mv, mp, mat, st, _ = payloads[r.rid]
            h = world.admit_scene(sid, mv, mp, mat, st)
            if h is None:
                still.append(r)
                continue
            active[h] = [r, r.steps]
            waits[r.rid] = now - r.arrival
        queue = still
        t = time.perf_counter()
        world.step(1)
        ticks.append(time.perf_counter() - t)
        done = []
        tick_bs = 0
        for h, a in active.items():
            tick_bs += payloads[a[0].rid][5]
            a[1] -= 1
            if a[1] == 0:
                done.append(h)
        body_steps += tick_bs
        tick_log.append((t - t0, tick_bs))
        # the finished scenes' results in one read
        results = world.scene_states(done, sum(len(payloads[active[h][0].rid][0]) for h in done))
        for h, (st, finite) in zip(done, results):
            r = active.pop(h)[0]
            if not finite:
                diverged.append(r.rid)
            if keep_states:
                states[r.rid] = st
            world.retire_scene(h)
            lat[r.rid] = time.perf_counter() - t0 - r.arrival
        if log and len(ticks) % 200 == 0:
            log(f"[serve] tick {len(ticks)}: {len(active)} active, {len(queue)} queued, {len(lat)}/{len(requests)} done")
    wall = time.perf_counter() - t0
    l = np.array([lat[r.rid] for r in requests])
    tk = np.array(ticks)
    # steady state: the ticks that start while requests are still arriving
    last = max(r.arrival for r in requests)
    win = [(ts, bs) for ts, bs in tick_log if ts <= last]
    steady = sum(bs for _, bs in win) / max(1e-9, (win[-1][0] + ticks[len(win) - 1]) if win else 1.0)
    return dict(requests=len(requests), wall_s=wall, body_steps=body_steps, value=body_steps / wall, steady_value=steady,
                p50_request_ms=1e3 * float(np.percentile(l, 50)), p99_request_ms=1e3 * float(np.percentile(l, 99)),
                p50_step_ms=1e3 * float(np.percentile(tk, 50)), p99_step_ms=1e3 * float(np.percentile(tk, 99)),
                ticks=len(ticks), mean_wait_ms=1e3 * float(np.mean(list(waits.values()))), states=states, diverged=diverged)
