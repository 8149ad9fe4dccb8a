"""Autotuner front end: irc_autotune (the library's tuner, include/irc_hip.h) on the caller's
device-resident batch, with its choice cached in a JSON file keyed by the workload shape so a later
process starts from it.

The library times irc_rho_delta_batch with each candidate setting of the screen's over-fetch depth
(k_over), persistent grid (waves_per_cu), smallest screened query group (tile: the queue
granularity) and deep-list threshold (large_ncand), and keeps the fastest whose outputs equal the
base setting's bit for bit (the settings only change how the exact result is reached).  Inside one
process, IRC_FLAG_AUTOTUNE applies the context's own choice without this module.
"""
import json
import os

from . import libirc

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "irc_hip_autotune.json")


def signature(dev, precision):
    """Cache key: precision, D, k and the sizes that set the screen's work balance."""
    F = dev["F"]
    n = int(dev["n_spikes"])
    S = len(dev["site_n_primary"])
    C = len(dev["lim_drift"]) - 1
    return f"p{precision}-D{int(F.shape[1])}-k{int(dev['k'])}-S{S}-C{C}-N{n >> 16}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, data):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(ctx, dev, out, precision=libirc.PREC_BF16_SCREEN, cache=DEFAULT_CACHE, force=False):
    """Return (options dict, report).  dev/out: device tensors as for Context.rho_delta_batch_device.
    The library's trial times are its own device time (HIP events around each call's kernels)."""
    key = signature(dev, precision)
    db = _load(cache) if cache else {}
    if not force and key in db:
        return dict(db[key]["options"]), {"cached": True, "key": key}
    choice, rep = ctx.autotune_device(dev, out, libirc.make_options(precision=precision))
    if cache:
        db[key] = {"options": choice, "ms": rep["ms_best"], "trials": rep["trials"]}
        try:
            _save(cache, db)
        except OSError:
            pass                      # a read-only home: the choice still applies to this run
    return choice, dict(rep, cached=False, key=key)
