"""The reference's on-disk inputs and outputs of this path.

* Tecplot ASCII structured multi-block solution files (INPUT/ic.tec, OUTPUT/field*.tec), the
  layout ORION's tec_read_structured_multiblock reads and writes (IO_Solution.f90:97-131):
  one VARIABLES line, then one ZONE per block with I, J, K node counts,
  DATAPACKING=BLOCK and VARLOCATION=([1-3]=NODAL,[4-n]=CELLCENTERED), the values in Fortran
  order (i fastest): x, y, z on the nodes, then the species densities rho1..rhoN and u, v, w,
  p on the cells.  One species here: rho = sum of the species densities.
* bc.txt (Read_BCfile, IO_BC.f90:136-414): per boundary face cell a line
  "b i j k f type" and, for the types that carry properties, a second line: the donor
  (bs is js ks fs d11 d12 d21 d22) for 101/201, (mach T p alpha beta rel_fac [massf...])
  for 405, (T0 p0 alpha beta rel_fac [massf...]) for 401 (stored in T and p), (p_back
  rel_fac) for 406 (p_back in p), where an angle may be the token 'normal' (parse_dir_tok,
  IO_BC.f90:443-456, stored as the largest double).  The number of records must match the blocks
  (Check_BC, IO_BC.f90:77-134).
"""
import gzip
import os
import re
import sys

import numpy as np

from . import abi, cases

NORMAL = sys.float_info.max          # the 'normal' direction token, like huge(1.0_R8)
# types without a property line (IO_BC.f90:107-119) and the ones this path runs
_NO_PROPS = {0, 200, 300, 400, 409}
_SUPPORTED = {0, 101, 201, 300, 400, 401, 405, 406}


def _open_text(path):
    return gzip.open(path, "rt") if str(path).endswith(".gz") else open(path)


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


# --------------------------------------------------------------------------- Tecplot
def _zone_header(text):
    h = {}
    for key in ("I", "J", "K"):
        m = re.search(r"\b%s\s*=\s*(\d+)" % key, text)
        if not m:
            raise ValueError("ZONE without %s" % key)
        h[key] = int(m.group(1))
    m = re.search(r"DATAPACKING\s*=\s*(\w+)", text, re.I)
    if m and m.group(1).upper() != "BLOCK":
        raise ValueError("only DATAPACKING=BLOCK is read")
    m = re.search(r"SOLUTIONTIME\s*=\s*([-+0-9.EeDd]+)", text, re.I)
    h["time"] = _num(m.group(1)) if m else 0.0
    m = re.search(r"VARLOCATION\s*=\s*\((.*?)\)\s*(,|$)", text, re.I)
    h["varloc"] = m.group(1) if m else ""
    m = re.search(r"\bT\s*=\s*\"?([^,\"]+)", text)
    h["title"] = m.group(1).strip() if m else ""
    return h


def _cell_centered(varloc, nvar):
    """Which variables (0-based) are cell-centred; default nodal."""
    cc = [False] * nvar
    for lo, hi, where in re.findall(r"\[(\d+)(?:-(\d+))?\]\s*=\s*(\w+)", varloc):
        a, b = int(lo), int(hi or lo)
        for v in range(a - 1, min(b, nvar)):
            cc[v] = where.upper().startswith("CELL")
    return cc


def read_tecplot(path):
    """Blocks of a structured multi-block Tecplot ASCII file: a list of dicts with dims
    (ni, nj, nk), nodes (nk+1, nj+1, ni+1, 3), prim (nk, nj, ni, 5) = (rho, u, v, w, p),
    time and the zone title."""
    with _open_text(path) as f:
        text = f.read()
    m = re.search(r"VARIABLES\s*=\s*(.*)", text, re.I)
    if not m:
        raise ValueError("no VARIABLES line")
    names = [n.strip().strip('"').lower() for n in re.findall(r'"[^"]*"|[^\s,"]+', m.group(1))]
    parts = re.split(r"^\s*ZONE\b", text, flags=re.I | re.M)[1:]
    if not parts:
        raise ValueError("no ZONE")
    species = [q for q, n in enumerate(names) if re.fullmatch(r"rho\d+", n)]
    need = {n: names.index(n) for n in ("x", "y", "z", "u", "v", "w", "p") if n in names}
    if len(need) != 7 or not species:
        raise ValueError("variables must include x y z rho1.. u v w p, found %s" % names)
    blocks = []
    for part in parts:
        lines = part.split("\n")
        q = 0
        header = []
        while q < len(lines):
            ln = lines[q].strip()
            if ln and re.match(r"^[-+]?(\d|\.\d)", ln):
                break
            header.append(ln)
            q += 1
        h = _zone_header(" ".join(header))
        data = np.array([_num(t) for t in " ".join(lines[q:]).replace(",", " ").split()])
        I, J, K = h["I"], h["J"], h["K"]
        ni, nj, nk = I - 1, J - 1, K - 1
        nn, nc = I * J * K, ni * nj * nk
        cc = _cell_centered(h["varloc"], len(names))
        sizes = [nc if c else nn for c in cc]
        if data.size < sum(sizes):
            raise ValueError("zone %s: %d values, %d expected" % (h["title"], data.size, sum(sizes)))
        arrays, o = [], 0
        for s, c in zip(sizes, cc):
            shape = (nk, nj, ni) if c else (K, J, I)
            arrays.append(data[o:o + s].reshape(shape))
            o += s
        for n in ("x", "y", "z"):
            if cc[need[n]]:
                raise ValueError("coordinates must be nodal")
        for n in ("u", "v", "w", "p"):
            if not cc[need[n]]:
                raise ValueError("flow variables must be cell-centred")
        nodes = np.stack([arrays[need["x"]], arrays[need["y"]], arrays[need["z"]]], axis=-1)
        prim = np.empty((nk, nj, ni, 5))
        prim[..., 0] = sum(arrays[s] for s in species)
        for v, n in enumerate(("u", "v", "w", "p"), start=1):
            prim[..., v] = arrays[need[n]]
        blocks.append({"dims": (ni, nj, nk), "nodes": np.ascontiguousarray(nodes), "prim": prim,
                       "time": h["time"], "title": h["title"]})
    return blocks


def write_tecplot(path, blocks, time=0.0):
    """Write blocks [(nodes (nk+1,nj+1,ni+1,3), prim (nk,nj,ni,5)), ...] in the layout of
    read_tecplot (one species), one value per line as the reference's files hold them (17 significant
    digits, so a round trip is exact)."""
    out = [' VARIABLES ="x" "y" "z"  "rho1" "u" "v" "w" "p"']
    for b, (nodes, prim) in enumerate(blocks, start=1):
        K, J, I = nodes.shape[:3]
        out.append(" ZONE  T = B%d-IG, I=%d, J=%d, K=%d, DATAPACKING=BLOCK, "
                   "VARLOCATION=([1-3]=NODAL,[4-8]=CELLCENTERED), SOLUTIONTIME=%.15E" % (b, I, J, K, time))
        cols = [nodes[..., 0], nodes[..., 1], nodes[..., 2]] + [prim[..., v] for v in range(5)]
        for c in cols:
            out.extend(" %.16E" % x for x in c.reshape(-1))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


# --------------------------------------------------------------------------- bc.txt
def _dir_tok(tok):
    return NORMAL if "normal" in tok.lower() else _num(tok)


def read_bc(path, dims_list=None):
    """bc.txt records in file order as an abi.BC_DTYPE array.  With the blocks' dims, the
    record count is checked as Check_BC does."""
    with _open_text(path) as f:
        lines = [ln for ln in f.read().split("\n") if ln.strip()]
    recs = []
    q = 0
    while q < len(lines):
        t = lines[q].replace(",", " ").split()
        q += 1
        if len(t) < 6:
            raise ValueError("bc record %d: '%s'" % (len(recs) + 1, lines[q - 1]))
        b, i, j, k, f, typ = (int(x) for x in t[:6])
        r = np.zeros((), dtype=abi.BC_DTYPE)
        r["b"], r["i"], r["j"], r["k"], r["f"], r["type"] = b, i, j, k, f, typ
        if typ not in _NO_PROPS:
            if q >= len(lines):
                raise ValueError("bc record %d (type %d): property line missing" % (len(recs) + 1, typ))
            p = lines[q].replace(",", " ").split()
            q += 1
            if typ in (101, 201):
                (r["bs"], r["is"], r["js"], r["ks"], r["fs"], r["d11"], r["d12"], r["d21"],
                 r["d22"]) = (int(x) for x in p[:9])
            elif typ == 405:
                r["mach"], r["T"], r["p"] = _num(p[0]), _num(p[1]), _num(p[2])
                r["alpha"], r["beta"] = _dir_tok(p[3]), _dir_tok(p[4])
                r["rel_fac"] = _num(p[5])
            elif typ == 401:
                r["T"], r["p"] = _num(p[0]), _num(p[1])
                r["alpha"], r["beta"] = _dir_tok(p[2]), _dir_tok(p[3])
                r["rel_fac"] = _num(p[4])
            elif typ == 406:
                r["p"], r["rel_fac"] = _num(p[0]), _num(p[1])
        if typ not in _SUPPORTED:
            raise ValueError("bc record %d (block %d face %d): type %d is not run by this path"
                             % (len(recs) + 1, b, f, typ))
        recs.append(r)
    arr = np.array(recs, dtype=abi.BC_DTYPE)
    if dims_list is not None:
        n = cases.expected_bc_count(dims_list)
        if len(arr) != n:
            raise ValueError("[ERROR] Boundary conditions number (%d) is different than the one of the "
                             "initial conditions (%d)" % (len(arr), n))
    return arr


def write_bc(path, bc):
    """bc.txt in the reference's layout."""
    out = []
    for r in bc:
        out.append("%8d%8d%8d%8d%8d%8d" % (r["b"], r["i"], r["j"], r["k"], r["f"], r["type"]))
        if r["type"] in (101, 201):
            out.append("%8d%8d%8d%8d%8d%8d%8d%8d%8d" % (r["bs"], r["is"], r["js"], r["ks"], r["fs"], r["d11"],
                                                         r["d12"], r["d21"], r["d22"]))
        elif r["type"] in (401, 405):
            ang = ["%15s" % "normal" if x >= 0.5 * NORMAL else "%15.6E" % x for x in (r["alpha"], r["beta"])]
            if r["type"] == 405:
                out.append("%15.6E,%15.6E,%15.6E,%s,%s,%15.6E,%15.6E," % (r["mach"], r["T"], r["p"], ang[0], ang[1],
                                                                         r["rel_fac"], 1.0))
            else:
                out.append("%15.6E,%15.6E,%s,%s,%15.6E,%15.6E," % (r["T"], r["p"], ang[0], ang[1], r["rel_fac"], 1.0))
        elif r["type"] == 406:
            out.append("%15.6E,%15.6E," % (r["p"], r["rel_fac"]))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def load_case(ic_path, bc_path, cfg, name=None):
    """A Case from an ic.tec and a bc.txt (the reference's INPUT directory)."""
    blocks = read_tecplot(ic_path)
    dims = [b["dims"] for b in blocks]
    bc = read_bc(bc_path, dims)
    cb = [cases.Block(b["dims"], b["nodes"], cases.padded_state(b["dims"], b["prim"])) for b in blocks]
    return cases.Case(name or os.path.basename(os.path.dirname(os.path.abspath(ic_path))), cfg, cb, bc,
                      t_end=None)
