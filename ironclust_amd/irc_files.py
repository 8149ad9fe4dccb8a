"""On-disk formats of irc2's detect and sort stages, so a real irc2 run can hand its features to the
HIP path and read the kNN back (host side; plain numpy file I/O).

  _fet_<iLoad>.irc, _fet2_<iLoad>.irc   raw column-major single arrays, no header: for every site
        in turn, trPc(:, :, spikes whose primary (secondary) site it is, ascending) -- nPc x
        nSites_spk values per spike, PC index fastest
        (detect_paged_save_ / save_paged_fet_site_, matlab/irc2.m:4583-4618)
  ccviSpk_site_load{iLoad}{iSite}   int64 global spike ids of those spikes: the load's local
        indices plus the count of the spikes of earlier loads (detect_merge_, matlab/irc2.m:4536-4577)
  load_fet_site_    one site's features over every load, masked by mlPc (matlab/irc2.m:4039-4135)
  get_mlPc_         the PC x site feature mask (matlab/irc2.m:3278-3312)
  _knn_<site>.irc   raw column-major int64 knn x n global spike ids, appended page by page
        (save_miKnn_site_, matlab/irc2.m:3355-3369, of miKnn_in = int64(...), 3670; read back as
        'int64' by load_miKnn_site_, 3395)

  sort.irc + sort_irc.mat   the sort stage's cache (sort_cache_, matlab/irc2.m:359-383): rho, delta,
        nneigh and ordrho written back to back by struct_save_bin_ (matlab/irc2.m:1733-1753), their
        class and size in the S_var struct saved beside them; struct_load_bin_ (1775-1805) reads it

`layout_from_fet_files` assembles the site-grouped column layout that irc_rho_delta_batch takes
(primary columns of a site from _fet_, then its secondary columns from _fet2_), which is exactly
what load_fet_site_(iFet = 1) and load_fet_site_(iFet = 2) return for each site.
"""
import os

import numpy as np


def write_bin_(f, arr, append=False):
    """fwrite(fid, mr, class(mr)) of a MATLAB array: raw bytes in column-major order
    (write_bin_, matlab/irc2.m:6891-6925).  f: path or an open binary file.  Returns dims, type."""
    a = np.asarray(arr)
    data = np.asfortranarray(a).tobytes(order="F")
    if hasattr(f, "write"):
        f.write(data)
    else:
        with open(f, "ab" if append else "wb") as fh:
            fh.write(data)
    return a.shape, a.dtype


def vi2cell_(viSite_spk, nSites):
    """cviSpk_site{iSite} = 1-based positions of the spikes of site iSite, ascending (a stable sort
    by site; sites < 1 dropped) (vi2cell_, matlab/irc2.m:4468-4499).  Returns a list of nSites
    int64 arrays (index iSite - 1)."""
    v = np.asarray(viSite_spk).astype(np.int64).ravel()
    order = np.argsort(v, kind="stable")
    out = [np.zeros(0, np.int64) for _ in range(nSites)]
    vs = v[order]
    for s in range(1, nSites + 1):
        lo, hi = np.searchsorted(vs, s, "left"), np.searchsorted(vs, s, "right")
        out[s - 1] = order[lo:hi] + 1
    return out


def _fet_path(prefix, iFet, iLoad):
    return f"{prefix}_fet{'' if iFet == 1 else '2'}_{iLoad}.irc"


def save_paged_fet_site_(fh, trFet_spk, viSite_spk, nSites):
    """Write trFet_spk (nPc x nSites_spk x n) site by site; return cviSpk_site (local, 1-based)
    (matlab/irc2.m:4612-4618)."""
    tr = np.asarray(trFet_spk, dtype=np.float32)
    cvi = vi2cell_(viSite_spk, nSites)
    for s in range(nSites):
        write_bin_(fh, tr[:, :, cvi[s] - 1])
    return cvi


def detect_paged_save_(prefix, iLoad, trPc_spk, viSite_spk, nSites, trPc2_spk=None, viSite2_spk=None):
    """Write _fet_<iLoad>.irc (and _fet2_ when secondary features exist) (matlab/irc2.m:4583-4609).
    Returns (cviSpk_site, cviSpk2_site, dimm_fet) with local 1-based spike indices."""
    with open(_fet_path(prefix, 1, iLoad), "wb") as fh:
        cvi1 = save_paged_fet_site_(fh, trPc_spk, viSite_spk, nSites)
    cvi2 = [np.zeros(0, np.int64) for _ in range(nSites)]
    if trPc2_spk is not None:
        with open(_fet_path(prefix, 2, iLoad), "wb") as fh:
            cvi2 = save_paged_fet_site_(fh, trPc2_spk, viSite2_spk, nSites)
    return cvi1, cvi2, tuple(np.asarray(trPc_spk).shape)


def merge_site_lists_(cviSpk_site_per_load, vnSpk_load):
    """ccviSpk_site_load: each load's local site lists offset by the spikes of earlier loads
    (cell_offset_ in detect_merge_, matlab/irc2.m:4541-4542, 4564)."""
    off = np.concatenate([[0], np.cumsum(np.asarray(vnSpk_load, dtype=np.int64))])
    return [[np.asarray(x, dtype=np.int64) + off[i] for x in cvi] for i, cvi in enumerate(cviSpk_site_per_load)]


def get_mlPc_(dimm_fet, nC_max, nPcPerChan=0, fMode_mlPc=2, nSites_fet=None):
    """mlPc (nPc x nSites_spk bool) and nFeatures (get_mlPc_, matlab/irc2.m:3278-3312)."""
    nPc_spk, nSites_spk = int(dimm_fet[0]), int(dimm_fet[1])
    npc = np.atleast_1d(np.asarray(nPcPerChan if nPcPerChan is not None else 0))
    if npc.size == 1:
        if int(npc[0]) == 0:
            if fMode_mlPc == 1:
                vn = np.arange(nPc_spk, 0, -1, dtype=np.int64)
            elif fMode_mlPc == 2:
                vn = np.full(nSites_spk, nPc_spk, dtype=np.int64)
            else:
                vn = np.full(int(nSites_fet), nPc_spk, dtype=np.int64)
            for i in range(len(vn) - 1, -1, -1):
                if vn.sum() <= nC_max:
                    break
                vn[i:] -= 1
                vn[vn < 0] = 0
        else:
            vn = np.full(nSites_spk, int(npc[0]), dtype=np.int64)
    else:
        vn = npc.astype(np.int64)
    ml = np.zeros((nPc_spk, nSites_spk), dtype=bool)
    for s in range(min(nSites_spk, len(vn))):
        ml[:int(vn[s]), s] = True
    return ml, int(ml.sum())


def load_fet_site_(prefix, iFet, iSite, ccviSpk_site_load, dimm_fet, mlPc=None):
    """One site's features over all loads (load_fet_site_, matlab/irc2.m:4039-4135, the load-all
    form): returns (mrFet nFeatures x n float32 when mlPc is given, else trPc nPc x nSites_spk x n;
    viSpk int64 global ids).  ccviSpk_site_load: per load, per site (0-based list index)."""
    nPc, nS = int(dimm_fet[0]), int(dimm_fet[1])
    per = nPc * nS
    parts, ids = [], []
    for iLoad, cvi in enumerate(ccviSpk_site_load, start=1):
        n1 = len(cvi[iSite - 1])
        if n1 == 0:
            continue
        off = sum(len(x) for x in cvi[:iSite - 1]) * per
        a = np.fromfile(_fet_path(prefix, iFet, iLoad), dtype=np.float32, count=n1 * per, offset=off * 4)
        parts.append(a.reshape((nPc, nS, n1), order="F"))
        ids.append(np.asarray(cvi[iSite - 1], dtype=np.int64))
    tr = np.concatenate(parts, axis=2) if parts else np.zeros((nPc, nS, 0), np.float32)
    vi = np.concatenate(ids) if ids else np.zeros(0, np.int64)
    if mlPc is None:
        return tr, vi
    m = np.asarray(mlPc, dtype=bool).ravel(order="F")
    return tr.reshape((per, -1), order="F")[m, :], vi


def save_miKnn_site_(prefix, iSite, miKnn1, fAppend=False):
    """Write (or append) a knn x n int64 block to _knn_<iSite>.irc (matlab/irc2.m:3355-3369): the
    reference's miKnn_in is int64 (matlab/irc2.m:3670), and fwrite writes its class."""
    path = f"{prefix}_knn_{iSite}.irc"
    a = np.asarray(miKnn1)
    if a.dtype != np.int64:
        if a.size and (a.min() < np.iinfo(np.int64).min or not np.issubdtype(a.dtype, np.integer)):
            raise TypeError("miKnn must hold integer spike ids")
        a = a.astype(np.int64)
    write_bin_(path, a, append=fAppend)
    return path


def load_miKnn_site_(prefix, iSite, knn):
    """Read _knn_<iSite>.irc back as knn x n int64 (load_bin_(..., 'int64', [knn, n]),
    matlab/irc2.m:3395)."""
    a = np.fromfile(f"{prefix}_knn_{iSite}.irc", dtype=np.int64)
    return a.reshape((knn, -1), order="F")


def layout_from_fet_files(prefix, ccviSpk_site_load, ccviSpk_site2_load, dimm_fet, mlPc, lim_drift, link,
                          n_spikes, k):
    """Site-grouped columns for irc_rho_delta_batch from the detect files: for each site, its
    primary columns (load_fet_site_ iFet = 1) then its secondary columns (iFet = 2); returns the
    dict irc2.sort_page_ / libirc.rho_delta_batch take (host arrays)."""
    nSites = len(ccviSpk_site_load[0])
    cols, spk, n_prim, n_all = [], [], [], []
    for s in range(1, nSites + 1):
        f1, v1 = load_fet_site_(prefix, 1, s, ccviSpk_site_load, dimm_fet, mlPc)
        have2 = ccviSpk_site2_load is not None and any(len(c[s - 1]) for c in ccviSpk_site2_load)
        if have2:
            f2, v2 = load_fet_site_(prefix, 2, s, ccviSpk_site2_load, dimm_fet, mlPc)
        else:
            f2, v2 = np.zeros((f1.shape[0], 0), np.float32), np.zeros(0, np.int64)
        cols += [f1.T, f2.T]
        spk += [v1, v2]
        n_prim.append(len(v1))
        n_all.append(len(v1) + len(v2))
    F = np.ascontiguousarray(np.concatenate(cols, axis=0), dtype=np.float32)
    return dict(F=F, col_spike=np.concatenate(spk).astype(np.int32),
                site_ptr=np.concatenate([[0], np.cumsum(n_all)]).astype(np.int64),
                site_n_primary=np.asarray(n_prim, dtype=np.int64),
                lim_drift=np.asarray(lim_drift, dtype=np.int64), link=np.asarray(link, dtype=np.uint8),
                n_spikes=int(n_spikes), k=int(k))


def save_knn_by_site(prefix, layout, knn, fAppend=False):
    """Per-site _knn_ files of a batch result: knn (n_spikes x k, or k x n_spikes) of 1-based spike
    ids; site s gets the columns of its primary spikes, ascending (the order sort_page_ appends
    them in).  Returns the paths."""
    K = np.asarray(knn)
    if K.shape[0] == layout["n_spikes"] and K.shape[1] != layout["n_spikes"]:
        K = K.T
    paths = []
    sp, npr, cs = layout["site_ptr"], layout["site_n_primary"], np.asarray(layout["col_spike"])
    for s in range(len(npr)):
        ids = cs[sp[s]:sp[s] + npr[s]].astype(np.int64)
        paths.append(save_miKnn_site_(prefix, s + 1, K[:, ids - 1], fAppend))
    return paths


# ---- the sort stage's cache: struct_save_bin_ / struct_load_bin_ ----------------------------------

_MATLAB_CLASS = {np.dtype(np.float32): "single", np.dtype(np.float64): "double", np.dtype(np.int64): "int64",
                 np.dtype(np.int32): "int32", np.dtype(np.uint32): "uint32", np.dtype(np.int16): "int16",
                 np.dtype(np.uint8): "uint8", np.dtype(np.bool_): "logical"}
_NUMPY_TYPE = {v: k for k, v in _MATLAB_CLASS.items()}


def _matlab_dims(a):
    """size() of the array as MATLAB reports it: vectors are columns (n x 1)."""
    return (a.shape[0], 1) if a.ndim == 1 else tuple(a.shape)


def struct_save_bin_(S, vcFile, csName_var):
    """[S0, S_var] = struct_save_bin_(S0, vcFile, csName_var) (matlab/irc2.m:1733-1753): the named
    fields are written back to back (fwrite of each class, column-major) into vcFile and cleared in the
    returned copy of S; S_var records vcFile, the names, classes (csType_var) and sizes (cDimm_var)."""
    S0 = dict(S)
    S_var = {"vcFile": vcFile, "csName_var": list(csName_var), "csType_var": [], "cDimm_var": []}
    with open(vcFile, "wb") as fh:
        for name in csName_var:
            a = np.asarray(S0[name])
            S0[name] = np.zeros((0, 0))
            if a.dtype not in _MATLAB_CLASS:
                raise TypeError(f"field {name}: no MATLAB class for {a.dtype}")
            write_bin_(fh, a)
            S_var["csType_var"].append(_MATLAB_CLASS[a.dtype])
            S_var["cDimm_var"].append(_matlab_dims(a))
    return S0, S_var


def struct_load_bin_(S_var, S=None, csVar_load=()):
    """S = struct_load_bin_(S_var, S, csVar_load) (matlab/irc2.m:1775-1805): each variable is read in
    turn with its class and size (fields not in csVar_load are skipped when it is given)."""
    S = dict(S or {})
    with open(S_var["vcFile"], "rb") as fh:
        for name, typ, dimm in zip(S_var["csName_var"], S_var["csType_var"], S_var["cDimm_var"]):
            dt = _NUMPY_TYPE[str(typ)]
            dimm = tuple(int(x) for x in np.ravel(dimm))
            n = int(np.prod(dimm))
            if csVar_load and not any(v in name for v in csVar_load):
                fh.seek(n * dt.itemsize, 1)
                continue
            a = np.fromfile(fh, dtype=dt, count=n)
            if a.size != n:
                raise IOError(f"{S_var['vcFile']}: {name} is truncated")
            S[name] = a.reshape(dimm, order="F")
    return S


def save_sort_(S_clu, vcFile_sort):
    """The save half of sort_cache_ (matlab/irc2.m:378-381): rho, delta, nneigh, ordrho go to
    <base>.irc through struct_save_bin_, everything else plus S_var to vcFile_sort (a MAT file;
    written as MAT v5, which MATLAB's load reads like the reference's v7.3).  vcFile_sort names the
    .mat file ('.../sort_irc.mat').  Returns (path of the .irc blob, S_var)."""
    from scipy.io import savemat
    vcFile_bin = vcFile_sort.replace("_irc.mat", ".irc")
    S_, S_var = struct_save_bin_(S_clu, vcFile_bin, ["rho", "delta", "nneigh", "ordrho"])
    def storable(v):   # what savemat writes as a MATLAB value (dicts become structs)
        if isinstance(v, dict):
            return all(storable(x) for x in v.values())
        return isinstance(v, (int, float, str, bool, np.integer, np.floating)) or (
            isinstance(v, np.ndarray) and v.dtype != object)

    mat = {k: ({kk: vv for kk, vv in v.items() if storable(vv)} if isinstance(v, dict) else v)
           for k, v in S_.items() if k not in S_var["csName_var"] and (isinstance(v, dict) or storable(v))}
    mat["S_var"] = {"vcFile": vcFile_bin,
                    "csName_var": np.array(S_var["csName_var"], dtype=object),
                    "csType_var": np.array(S_var["csType_var"], dtype=object),
                    "cDimm_var": np.array([np.array(d, dtype=np.float64) for d in S_var["cDimm_var"]], dtype=object)}
    savemat(vcFile_sort, mat)
    return vcFile_bin, S_var


def load_sort_(vcFile_sort):
    """The load half of sort_cache_ (matlab/irc2.m:366-368): S_clu = load(vcFile_sort), then
    struct_load_bin_(S_clu.S_var, S_clu)."""
    from scipy.io import loadmat
    m = loadmat(vcFile_sort, squeeze_me=True, simplify_cells=True)
    sv = m.pop("S_var")
    S_var = {"vcFile": str(sv["vcFile"]), "csName_var": [str(x) for x in np.ravel(sv["csName_var"])],
             "csType_var": [str(x) for x in np.ravel(sv["csType_var"])],
             "cDimm_var": [np.ravel(x).astype(np.int64) for x in sv["cDimm_var"]]}
    S = {k: v for k, v in m.items() if not k.startswith("__")}
    return struct_load_bin_(S_var, S)
