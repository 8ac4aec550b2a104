"""CPU oracle for the boxcar S/N of a dynamic spectrum (test infrastructure only).

Plain-Python restatement of the reference's ``_calculate_frb_snr``
(src/bindpy/frb_snr.cpp:144-398; helpers :12-140), for small inputs.  Only
tests/ import it.  Pinned in tests/test_snr_cpu.py by closed forms: a
rectangular pulse on a +-1 alternating baseline (exact mean 0 / std 1 noise,
S/N = A*sqrt(w)), the width list's closed form, the clipped statistics of
hand-made sets, and the empty / fallback returns.
"""
from __future__ import annotations

import math


def _cdiv(a, b):
    """C integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def median(v):
    """frb_snr.cpp:32-45 -- mean of the two middle values for an even count."""
    s = sorted(v)
    n = len(s)
    if n == 0:
        return 0.0
    if n % 2:
        return s[n // 2]
    return 0.5 * (s[n // 2 - 1] + s[n // 2])


def mean(v):
    """frb_snr.cpp:12-18 (sequential double sum)."""
    if not v:
        return 0.0
    acc = 0.0
    for x in v:
        acc += x
    return acc / len(v)


def pstd(v, mu):
    """frb_snr.cpp:20-30 (population std about mu)."""
    if not v:
        return 0.0
    acc = 0.0
    for x in v:
        acc += (x - mu) * (x - mu)
    return math.sqrt(acc / len(v))


def robust(data, sigma, rounds=3):
    """frb_snr.cpp:47-100: sigma clipping about the median with 1.4826*MAD."""
    fin = [x for x in data if math.isfinite(x)]
    if not fin:
        return 0.0, 1.0
    cur = list(fin)
    for _ in range(rounds):
        med = median(cur)
        mad = median([abs(x - med) for x in cur])
        sd = 1.4826 * mad if mad > 0 else pstd(cur, mean(cur))
        if sd <= 0:
            break
        nxt = [x for x in cur if abs(x - med) <= sigma * sd]
        if len(nxt) == len(cur):
            break
        if len(nxt) < max(10, int(0.2 * len(cur))):
            break
        cur = nxt
    mu = mean(cur)
    sd = pstd(cur, mu)
    if sd <= 0:
        sd = pstd(fin, mean(fin))
    if sd <= 0:
        sd = 1.0
    return mu, sd


def widths(wmax):
    """frb_snr.cpp:102-140."""
    if wmax <= 1:
        return [1]
    if wmax <= 16:
        return list(range(1, wmax + 1))
    lo, hi = math.log10(17.0), math.log10(float(wmax))
    extra = set()
    for i in range(32):
        v = 10.0 ** (lo + (i / 31.0) * (hi - lo))
        r = int(math.floor(v + 0.5))          # llround of a positive value
        if 17 <= r <= wmax:
            extra.add(r)
    return list(range(1, 17)) + sorted(extra)


def frb_snr(spec, noise_range=None, threshold_sigma=5.0, toa_sample_idx=None,
            fitting_window_samples=None, tsamp=None, target_time_us=256.0):
    """frb_snr.cpp:144-398: returns (snr, width, peak, (mean, std, info))."""
    nt0 = len(spec)
    nf = len(spec[0]) if nt0 else 0
    if nt0 <= 0 or nf <= 0:
        return -1.0, 0, 0, (0.0, 1.0, {"fit_converged": False})
    ts = []
    for row in spec:
        acc = 0.0
        for v in row:
            acc += float(v)
        ts.append(acc)
    ds = 1
    if tsamp is not None and target_time_us is not None:
        if tsamp > 0:
            q = target_time_us * 1e-6 / tsamp
            ds = int(math.floor(abs(q) + 0.5)) * (1 if q >= 0 else -1)
        ds = max(ds, 1)
    if ds > 1:
        ng = nt0 // ds
        if ng >= 1:
            ts2 = []
            for g in range(ng):
                acc = 0.0
                for i in range(ds):
                    acc += ts[g * ds + i]
                ts2.append(acc)
            ts = ts2
        else:
            ds = 1
    n = len(ts)
    if fitting_window_samples is None:
        wmax = max(8, min((n * ds) // 4, 512))
    else:
        wmax = max(4, min(int(fitting_window_samples), (n * ds) // 2))
    wmax = max(1, min(wmax, n * ds))
    wmax_ds = min(int(math.ceil(wmax / ds)), n)
    s0, s1 = 0, n
    if toa_sample_idx is not None:
        rad = max(wmax_ds, int(math.ceil(100.0 / ds)))
        c = _cdiv(int(toa_sample_idx), ds)
        s0, s1 = max(0, c - rad), min(n, c + rad)
    noise = []
    if noise_range is not None and len(noise_range) > 0:
        for item in noise_range:
            if len(item) != 2:
                continue
            a, b = int(item[0]), int(item[1])
            a_ds = max(0, _cdiv(a, ds))
            b_ds = min(n, int(math.ceil(b / ds)))
            if b_ds > a_ds:
                noise.extend(ts[a_ds:b_ds])
        if not noise:
            noise = list(ts)
    else:
        if toa_sample_idx is not None and (s0 > 0 or s1 < n):
            noise = ts[:s0] + ts[s1:]
        if len(noise) < 10:
            noise = list(ts)
    mu, sd = robust(noise, threshold_sigma)
    # window sums as differences of running sums (frb_snr.cpp:290-296,321)
    run = [0.0]
    for v in ts:
        run.append(run[-1] + v)
    peak = max(range(n), key=lambda i: (ts[i], -i))
    best, best_sum, best_w, best_c = -math.inf, math.nan, 1, peak
    for w in widths(wmax_ds):
        if w >= n:
            break
        found, lb, lc, ls = False, -math.inf, best_c, 0.0
        for i in range(n + 1 - w):
            c = w // 2 + i
            if toa_sample_idx is not None and (c < s0 or c >= s1):
                continue
            ssum = run[i + w] - run[i]
            snr = (ssum - mu * w) / (sd * math.sqrt(w)) if sd > 0 else -math.inf
            if not found or snr > lb:
                found, lb, lc, ls = True, snr, c, ssum
        if found and lb > best:
            best, best_w, best_c, best_sum = lb, w, lc, ls
    if not math.isfinite(best_sum):
        m2, s2 = robust(ts, threshold_sigma)
        snr = (ts[peak] - m2) / s2 if s2 > 0 else -1.0
        return snr, max(1, ds), min(nt0 - 1, peak * ds + ds // 2), \
            (m2, s2, {"fit_converged": False, "method": "fallback"})
    rl, rr = max(0, best_c - 2 * best_w), min(n, best_c + 2 * best_w + 1)
    rest = [ts[i] for i in range(n) if i < rl or i >= rr]
    if len(rest) >= max(10, int(0.1 * n)):
        mu, sd = robust(rest, threshold_sigma)
    snr = (best_sum - mu * best_w) / (sd * math.sqrt(best_w)) if sd > 0 else -1.0
    info = {"fit_converged": True, "method": "boxcar", "width_samples": best_w * ds,
            "search_start": s0, "search_end": s1, "downsample": ds}
    return snr, max(1, best_w * ds), min(nt0 - 1, best_c * ds + ds // 2), (mu, sd, info)

