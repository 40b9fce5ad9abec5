"""Open-loop driver for the serving loop: camera streams with Poisson arrivals.

Frames of each stream come from synth.stream_frames (each frame the previous
one moved by a small rotation and shift, plus fresh noise) and are replayed
forwards then backwards (0, 1, .., n-1, n-2, .., 1, 0, 1, ..), so that every
request's previous frame on its stream is a neighbour in the sequence.
Arrival times are drawn per stream (exponential inter-arrival at `fps`) and
submitted at their scheduled time; latency is completion time minus scheduled
arrival on the library clock.
"""
import time

import numpy as np

from . import ORB_E_ADMISSION, OrbError, Server, synth

MIXED_16 = [(640, 480, 1000)] * 4 + [(752, 480, 1000)] * 4 + [(1241, 376, 2000)] * 4 + [(1280, 720, 2000)] * 4


def make_stream_frames(streams, n_frames, seed=100):
    return [synth.stream_frames(w, h, seed + i, n_frames) for i, (w, h, _) in enumerate(streams)]


def sequence_index(k, n):
    """Frame of request k on a stream of n frames, replayed back and forth."""
    if n <= 1:
        return 0
    j = k % (2 * n - 2)
    return j if j < n else 2 * n - 2 - j


def poisson_schedule(n_streams, fps, duration, seed):
    rng = np.random.default_rng(seed)
    events = []
    for s in range(n_streams):
        t, k = rng.exponential(1.0 / fps), 0
        while t < duration:
            events.append((t, s, k))
            k += 1
            t += rng.exponential(1.0 / fps)
    events.sort()
    return events


def run(server, frames, fps=30.0, duration=3.0, seed=7, keep=None):
    """Replays a Poisson schedule; returns per-request records (and the
    fetched results of requests whose (stream, index) is in `keep`)."""
    events = poisson_schedule(len(frames), fps, duration, seed)
    records, kept, by_ticket = [], {}, {}
    refused = 0
    t0 = server.clock() + 0.05
    i = 0
    while i < len(events) or len(records) + refused < len(events):
        now = server.clock()
        while i < len(events) and t0 + events[i][0] <= now:
            t_rel, s, k = events[i]
            try:
                ticket = server.submit(s, frames[s][sequence_index(k, len(frames[s]))], t0 + t_rel)
                by_ticket[ticket] = (s, k)
            except OrbError as e:
                if e.rc != ORB_E_ADMISSION:
                    raise
                refused += 1
            i += 1
        for rec in server.poll():
            s, k = by_ticket.pop(rec["ticket"])
            rec.update(stream=s, index=k)
            records.append(rec)
            if keep is not None and (s, k) in keep:
                kept[(s, k)] = server.fetch(rec["ticket"])
            else:
                server.release(rec["ticket"])
        if i < len(events):
            wait = t0 + events[i][0] - server.clock()
            if wait > 0.0005:
                time.sleep(min(wait - 0.0003, 0.002))
        elif len(records) + refused < len(events):
            time.sleep(0.0002)
    for r in records:
        r["refused_in_run"] = refused
    return records, kept


def summarize(records):
    lat = np.array([r["t_done"] - r["t_arrival"] for r in records]) * 1e3
    t_first = min(r["t_arrival"] for r in records)
    t_last = max(r["t_done"] for r in records)
    return {"requests": len(records), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(lat.max()), 3),
            "frames_per_s": round(len(records) / (t_last - t_first), 1),
            "errors": int(sum(r["status"] != 0 for r in records)),
            "refused": int(records[0].get("refused_in_run", 0)) if records else 0}
