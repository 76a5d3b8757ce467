"""Device basis pack: the flat arrays handed to ``ace_ctx_create``.

Built once per basis on the CPU from the shared tables
(:func:`acehip.tables.basis_tables`).  Holds the radial recursion, the
pooled ``m >= 0`` table, the AA / A2B tables for the basis-level calls, and
the run programs of the energy and of the gradient polynomials together with
their task -> wave assignments for every supported tile shape.
"""
from __future__ import annotations

import numpy as np

from . import programs as P

WAVE_CONFIGS = (4, 8, 16)     # waves per workgroup of the product-DAG kernel
ENERGY_PARTS = 16             # energy polynomial split into this many tasks
REC_LEAVES = 4                # leaves per 64-byte program record
MAX_TASK_COST = 1200.0        # gradient polynomials above this are split (tasks fit a quarter-tile wave)
MAX_FEATURES = 160            # order <= 2 invariants the surrogate kernels take


class Pack:
    def __init__(self, T: dict):
        self.T = T
        self.prop = 0 if T["prop"] == "Invariant" else 1
        aspec = np.asarray(T["aspec"], dtype=np.int64)
        self.pool, kk, cj, sgn = P.pool_map(aspec)
        self.aspec_fid = (kk | (cj * P.DEV_CONJ)).astype(np.int32)
        self.aspec_sgn = sgn.astype(np.float64)
        self.nspecies = 0 if T["categories"] is None else len(T["categories"])
        orders = np.asarray(T["orders"], dtype=np.int64)
        iAA2iA = np.asarray(T["iAA2iA"], dtype=np.int64)
        if iAA2iA.shape[1] > P.MAX_ORDER:
            raise ValueError("correlation order above 8 is not supported")
        is_m0 = self.pool[:, 3] == 0
        # A2B CSR (rows ascending, columns ascending within a row)
        a2b = T["a2b"]
        rows_ = np.asarray(a2b["rows"], dtype=np.int64)
        cols_ = np.asarray(a2b["cols"], dtype=np.int64)
        vals_ = np.asarray(a2b["vals"])
        order = np.lexsort((cols_, rows_))
        rows_, cols_, vals_ = rows_[order], cols_[order], vals_[order]
        # EuclideanVector: class-split programs when the A2B map allows it
        self.classed = 0
        self.row_class = P.row_classes(orders, iAA2iA, aspec)
        if self.prop == 1 and P.a2b_is_classed(cols_, vals_, self.row_class):
            self.classed = 1
        poly, _ = P.energy_polynomial(orders, iAA2iA, kk, cj, sgn)
        poly = P.merge_conjugates(poly, is_m0)
        self.poly = poly
        G = P.gradient_polynomials(poly, len(self.pool), is_m0)
        prog = P.RunProgram()
        if self.classed:
            # energy: one task group per class Q_M, output slot -1 - class
            self.energy_tasks = []
            for c, q in enumerate(P.split_by_class(poly, self.row_class)):
                if q:
                    self.energy_tasks += [prog.add_task(part, -1 - c) for part in P.split_poly(q, ENERGY_PARTS)]
        else:
            self.energy_tasks = [prog.add_task(part, -1) for part in P.split_poly(poly, ENERGY_PARTS)]
        # gradient tasks: split large polynomials over several output slots
        # (and, class-split programs, by class: slot class in gslot_class)
        gslots = [[] for _ in range(len(self.pool))]
        gclass = []
        self.grad_tasks = []
        nslot = 0
        for k, g in enumerate(G):
            if not g:
                continue
            for c, gc in (enumerate(P.split_by_class(g, self.row_class)) if self.classed else [(0, g)]):
                if not gc:
                    continue
                cost = sum(1.0 for m in gc) + 0.0
                nparts = max(1, int(np.ceil(cost / MAX_TASK_COST)))
                parts = P.split_poly(gc, nparts) if nparts > 1 else [gc]
                for part in parts:
                    t = prog.add_task(part, nslot)
                    self.grad_tasks.append(t)
                    gslots[k].append(nslot)
                    gclass.append(c)
                    nslot += 1
        self.ngslots = nslot
        self.gslot_class = np.asarray(gclass if self.classed else [], dtype=np.int32)
        self.gslot_ptr = np.zeros(len(self.pool) + 1, dtype=np.int32)
        self.gslot_ptr[1:] = np.cumsum([len(s) for s in gslots])
        self.gslot_list = np.asarray([s for ss in gslots for s in ss], dtype=np.int32)
        self.prog = prog
        # flatten programs
        self.runs = np.asarray(prog.runs, dtype=np.int32).reshape(-1, P.RUN_W)
        self.leaf_fid = np.asarray(prog.leaf_fid, dtype=np.int32)
        tasks3 = np.asarray(prog.tasks, dtype=np.int32).reshape(-1, 3)
        self._build_records(tasks3, records=iAA2iA.shape[1] <= 4 and (self.prop == 0 or self.classed == 1))
        lt_ptr, rows, scales, conjs = [0], [], [], []
        for terms in prog.leaf_terms:
            for (r, s, c) in terms:
                rows.append(r)
                scales.append(s)
                conjs.append(c)
            lt_ptr.append(len(rows))
        tt_ptr = [len(rows)]
        for terms in prog.task_terms:
            for (r, s, c) in terms:
                rows.append(r)
                scales.append(s)
                conjs.append(c)
            tt_ptr.append(len(rows))
        # task terms follow the leaf terms in the same arrays
        self.leaf_term_ptr = np.asarray(lt_ptr, dtype=np.int32)
        self.task_term_ptr = np.asarray(tt_ptr, dtype=np.int32)
        self.term_row = np.asarray(rows, dtype=np.int32)
        self.term_scale = np.asarray(scales, dtype=np.float64)
        self.term_conj = np.asarray(conjs, dtype=np.int32)
        # wave assignments
        costs = [prog.cost(t) for t in range(len(prog.tasks))]
        self.wave_cfg_W = np.asarray(WAVE_CONFIGS, dtype=np.int32)
        wtp, wtl, wep, wel = [], [], [], []
        self.wave_load = {}
        for W in WAVE_CONFIGS:
            lists, load = P.lpt_assign(costs, W)
            wtp.extend(np.concatenate([[0], np.cumsum([len(l) for l in lists])]).tolist())
            for l in lists:
                wtl.extend(l)
            ecost = [costs[t] for t in self.energy_tasks]
            elists, eload = P.lpt_assign(ecost, W)
            elists = [[self.energy_tasks[i] for i in l] for l in elists]
            wep.extend(np.concatenate([[0], np.cumsum([len(l) for l in elists])]).tolist())
            for l in elists:
                wel.extend(l)
            self.wave_load[W] = (max(load), sum(load), max(eload), sum(eload))
        self.wave_task_ptr = np.asarray(wtp, dtype=np.int32)
        self.wave_task_list = np.asarray(wtl, dtype=np.int32)
        self.wave_etask_ptr = np.asarray(wep, dtype=np.int32)
        self.wave_etask_list = np.asarray(wel, dtype=np.int32)
        nB, nAA = a2b["shape"]
        self.nB, self.nAA = int(nB), int(nAA)
        rp = np.zeros(self.nB + 1, dtype=np.int64)
        np.add.at(rp, rows_ + 1, 1)
        self.a2b_rowptr = np.cumsum(rp).astype(np.int32)
        self.a2b_cols = cols_.astype(np.int32)
        if self.prop == 0:
            self.a2b_vals = np.asarray(vals_, dtype=np.float64)
        else:
            v = np.asarray(vals_, dtype=np.complex128).reshape(-1, 3)
            self.a2b_vals = np.stack([v.real, v.imag], axis=-1).reshape(-1)
        self.orders = orders.astype(np.int32)
        self.maxorder = int(iAA2iA.shape[1])
        self.iAA2iA = iAA2iA.astype(np.int32)
        if self.prop == 0:
            self._build_features(rows_, cols_, np.asarray(vals_, dtype=np.float64), kk, cj, sgn)
        else:
            self._empty_features()

    def _empty_features(self):
        self.feat_rows = np.zeros(0, dtype=np.int32)
        self.const_rows = np.zeros(0, dtype=np.int32)
        self.feat_ptr = np.zeros(1, dtype=np.int32)
        self.feat_f = np.zeros((0, 2), dtype=np.int32)
        self.feat_coef = np.zeros(0)
        self.fgrad_ptr = np.zeros(len(self.pool) + 1, dtype=np.int32)
        self.fgrad_feat = np.zeros(0, dtype=np.int32)
        self.fgrad_f = np.zeros(0, dtype=np.int32)
        self.fgrad_coef = np.zeros(0)

    def _build_features(self, rows, cols, vals, kk, cj, sgn):
        """Order 1 and 2 invariants (the surrogate's inputs): the B rows whose AA
        columns all have order 1 or 2 (every B row has a single order).  The
        order-0 row is the constant already carried by the energy constant,
        so it is not a feature.
        Feature f = sum_terms coef * Re(A_f1 A_f2) with f1/f2 device factor ids
        (-1: absent), coef = U * sign.  Gradient entries, grouped by target
        pooled index k, give G[k] += w_f * coef * value(other) with
        value(-1) = 1 (see programs.py for the definition of G)."""
        o = self.orders[cols]
        maxo = np.zeros(self.nB, dtype=np.int64)
        np.maximum.at(maxo, rows, o)
        feat_rows = np.nonzero((maxo >= 1) & (maxo <= 2))[0]
        if len(feat_rows) > MAX_FEATURES:
            self._empty_features()  # more invariants than the surrogate kernels take
            return
        self.feat_rows = feat_rows.astype(np.int32)
        self.const_rows = np.nonzero(maxo == 0)[0].astype(np.int32)
        fptr, ff, fc = [0], [], []
        grad = [[] for _ in range(len(self.pool))]
        dev = lambda a: int(kk[a] | (cj[a] * P.DEV_CONJ))  # noqa: E731
        for fi, b in enumerate(feat_rows):
            for z in np.nonzero(rows == b)[0]:
                col = int(cols[z])
                od = int(self.orders[col])
                idx = self.iAA2iA[col, :od]
                s = float(np.prod(sgn[idx])) if od else 1.0
                c = float(vals[z]) * s
                f = [dev(a) for a in idx] + [-1] * (2 - od)
                ff.append(f)
                fc.append(c)
                # gradient entries: d/dA+[k] directly, d/dconj(A+[k]) conjugated
                if od == 1:
                    grad[f[0] & 0x7fff].append((fi, -1, c))
                elif od == 2:
                    for a_, b_ in ((0, 1), (1, 0)):
                        fa, fb = f[a_], f[b_]
                        k = fa & 0x7fff
                        if fa & P.DEV_CONJ:
                            # conj(d/dconj A+[k]) = conj(A_fb): flip fb's conjugation
                            other = fb ^ P.DEV_CONJ if not self._is_m0(fb) else fb
                        else:
                            other = fb
                        grad[k].append((fi, other, c))
            fptr.append(len(ff))
        self.feat_ptr = np.asarray(fptr, dtype=np.int32)
        self.feat_f = np.asarray(ff, dtype=np.int32).reshape(-1, 2)
        self.feat_coef = np.asarray(fc, dtype=np.float64)
        gp, gf, go, gc = [0], [], [], []
        for k in range(len(self.pool)):
            for (fi, other, c) in grad[k]:
                gf.append(fi)
                go.append(other)
                gc.append(c)
            gp.append(len(gf))
        self.fgrad_ptr = np.asarray(gp, dtype=np.int32)
        self.fgrad_feat = np.asarray(gf, dtype=np.int32)
        self.fgrad_f = np.asarray(go, dtype=np.int32)
        self.fgrad_coef = np.asarray(gc, dtype=np.float64)

    def _is_m0(self, f):
        return self.pool[f & 0x7fff, 3] == 0

    def _build_records(self, tasks3, records=True):
        """64-byte records: 8 header dwords + REC_LEAVES coefficients.

        dw0 = depth | jstart << 2 | nleaf << 5; dw1 = f1 | f2 << 16;
        dw2 = f3; dw3.. = leaf factor ids, two per dword.  A run with more
        than REC_LEAVES leaves continues in further records with the same
        prefix (jstart = depth + 1).  ``rec_leaf[r, i]`` is the leaf whose
        coefficient fills slot i (-1: empty).  Records hold three prefix
        factors: bases above correlation order 4 get none (the run
        interpreter evaluates them)."""
        K = REC_LEAVES
        hdr, rleaf = [], []
        rec_of_task = []
        for (r0, r1, slot) in tasks3:
            b = len(hdr)
            for r in range(r0, r1 if records else r0):
                d, j, nl, lo, f1, f2, f3 = (int(x) for x in self.runs[r][:7])
                for q0 in range(0, max(nl, 1), K):
                    n = min(K, nl - q0)
                    jj = j if q0 == 0 else d + 1
                    h = [0] * 8
                    h[0] = d | (jj << 2) | (n << 5)
                    h[1] = f1 | (f2 << 16)
                    h[2] = f3
                    ids = [int(self.leaf_fid[lo + q0 + i]) if i < n else 0 for i in range(K)]
                    for i in range(0, K, 2):
                        h[3 + i // 2] = ids[i] | (ids[i + 1] << 16)
                    hdr.append(h)
                    rleaf.append([lo + q0 + i if i < n else -1 for i in range(K)])
            rec_of_task.append((b, len(hdr)))
        self.rec_hdr = np.asarray(hdr, dtype=np.uint32).reshape(-1, 8)
        self.rec_leaf = np.asarray(rleaf, dtype=np.int32).reshape(-1, K)
        self.tasks = np.concatenate([tasks3, np.asarray(rec_of_task, dtype=np.int32).reshape(-1, 2)], axis=1)

    # -- statistics used by the roofline accounting
    def work_counts(self, forces: bool):
        """Executed leaf and run counts per environment of the product-DAG
        kernel (energy tasks, plus gradient tasks with forces)."""
        tasks = list(self.energy_tasks) + (list(self.grad_tasks) if forces else [])
        leaves = runs = 0
        newf = 0
        for t in tasks:
            r0, r1 = self.tasks[t][0], self.tasks[t][1]
            rr = self.runs[r0:r1]
            leaves += int(rr[:, 2].sum())
            runs += len(rr)
            newf += int(np.sum(np.maximum(rr[:, 0] - rr[:, 1] + 1, 0)))
        recs = sum(int(self.tasks[t][4] - self.tasks[t][3]) for t in tasks)
        return {"leaves": leaves, "runs": runs, "prefix_mults": newf, "records": recs}
