#!/usr/bin/env python3
"""Handler emitter: the back end of cdvmh, producing gfx950 HIP.

Input: a loop description file (<program>.loops, a Python literal) that lists,
for each parallel loop of a converted program, what cdvmh's front end knows
about it: loop variables and constant steps, the handler parameters (DVM
arrays passed as headers, scalars passed by pointer), private variables,
reduction clauses, the ACROSS clause and the loop body text.

Output, per program:
  <program>_host.c   host handlers (C + OpenMP), the shape of cdvmh's
                     genHostHandler (cdvmh/src/converter_main.cpp:1503-1859):
                     red_init/red_post per thread, "omp for" on the first
                     independent axis, the pipelined handshake when every axis
                     carries an ACROSS dependency.
  <program>_hip.hip  GPU handlers + kernels.  The handler makes the same ABI
                     calls as cdvmh's genCudaHandler (converter_main.cpp:
                     2132-2401) / genAcrossCudaHandler (across.cpp:2788-2894),
                     asks the runtime for its tuned configuration
                     (dvmh_amd_loop_get_tuning_C) and launches kernels designed
                     for wave64/LDS:
                       PARALLEL   one launch; x <- innermost axis, y, z next;
                                  DEPTH consecutive iterations of the slowest
                                  mapped axis per thread (an unguarded copy for
                                  threads whose iterations are all inside the
                                  loop); XCD-aware tile order; in-kernel
                                  wave+LDS reduction, one partial per workgroup.
                       MARCH      variant of PARALLEL for loops that read some
                                  arrays as a radius-1 stencil: a workgroup
                                  marches a tile along the outermost axis with a
                                  rolling three-plane LDS window per stencil
                                  array, the next plane's loads in flight while
                                  the current plane is computed.
                       LINE       ACROSS with one dependent axis: one launch,
                                  threads over the independent axes, each
                                  thread walks its line in dependency order.
                       WAVEFRONT  ACROSS with >= 2 dependent axes: one launch
                                  per hyperplane of the dependent axes; arrays
                                  may be in the diagonal layout.
The loop body is inserted verbatim; every schedule respects the dependency
order, so results are bit-identical to the sequential host order.
"""
import ast
import itertools
import os
import re
import sys

RT = {'char': 'rt_CHAR', 'int': 'rt_INT', 'long': 'rt_LONG', 'float': 'rt_FLOAT', 'double': 'rt_DOUBLE',
      'long long': 'rt_LLONG', 'unsigned': 'rt_UINT', 'unsigned long': 'rt_ULONG'}
RED_FUNC = {'sum': 'rf_SUM', 'prod': 'rf_PROD', 'max': 'rf_MAX', 'min': 'rf_MIN', 'and': 'rf_AND', 'or': 'rf_OR',
            'xor': 'rf_XOR', 'maxloc': 'rf_MAXLOC', 'minloc': 'rf_MINLOC'}
RED_OP = {'sum': 'OpSum', 'prod': 'OpProd', 'max': 'OpMax', 'min': 'OpMin', 'and': 'OpAnd', 'or': 'OpOr', 'xor': 'OpXor'}
OMP_RED = {'sum': '+', 'prod': '*', 'max': 'max', 'min': 'min', 'and': '&', 'or': '|', 'xor': '^'}
IND = '    '


def indent(text, n):
    pad = IND * n
    return ''.join(pad + ln if ln.strip() else ln for ln in text.splitlines(True))


class Loop:
    def __init__(self, d, prog):
        self.prog = prog
        self.name = d['name']
        self.vars = d['vars']                      # [(name, ctype)] outermost first
        self.rank = len(self.vars)
        self.steps = d.get('steps', ['1'] * self.rank)   # constant steps as text, None = runtime
        self.params = d['params']                  # [('array', name, ctype, rank) | ('scalar', name, ctype)
                                                   #  | ('carray', name, ctype, [extents])]
        self.host_only = d.get('host_only', False)  # a loop outside regions: no GPU handler
        # a region's statement outside parallel loops: a loop of rank 0 whose
        # handlers run the statement once (cdvmh's sequential part)
        self.sequential = d.get('sequential', False)
        self.privates = d.get('privates', [])      # [(decl text)]
        self.reds = d.get('reductions', [])        # [(func, var, ctype)] or [(func, var, ctype, loc, loctype)]
        # reduction arrays: variable -> element count (a C array of that many
        # elements, each reduced on its own; value reductions only)
        self.red_lengths = dict(d.get('reduction_lengths', {}))
        for red in self.reds:
            if self.rlen(red) > 1 and len(red) > 3:
                raise ValueError(f'{self.name}: {red[0]} over the array {red[1]} is not supported')
        self.across = d.get('across')              # None | {'array': name, 'widths': [(lo, hi), ...]}
        self.body = d['body'].strip('\n')
        self.cuda_block = d.get('cuda_block')      # optional default block
        self.batch = d.get('batch', False)         # also emit a batch handler
        self.arrays = [p for p in self.params if p[0] == 'array']
        self.scalars = [p for p in self.params if p[0] == 'scalar']
        # arrays read as a radius-1 stencil around the loop point: candidates for
        # the LDS march variant (checked against the body below)
        self.stencil = [a for a in d.get('stencil', []) if self._is_stencil(a)]
        if self.batch and any(self.rlen(red) > 1 for red in self.reds):
            raise ValueError(f'{self.name}: a batch handler takes scalar reductions only')

    def rlen(self, red):
        """Elements of a reduction variable (1 for a scalar)."""
        return int(self.red_lengths.get(red[1], 1))

    def _is_stencil(self, name):
        """True when the body only reads `name` and every access is
        name[v0 +- 0/1][v1 +- 0/1]... with the loop variables in order."""
        if self.across or self.rank not in (2, 3):
            return False
        arr = [p for p in self.arrays if p[1] == name]
        if not arr or arr[0][3] != self.rank:
            return False
        if re.search(r'\b' + name + r'\s*(\[[^\[\]]*\]\s*)+(=(?!=)|\+=|-=|\*=|/=|\+\+|--)', self.body):
            return False
        if re.search(r'(\+\+|--)\s*' + name + r'\b', self.body):
            return False
        for m in re.finditer(r'\b' + name + r'\s*((?:\[[^\[\]]*\]\s*)+)', self.body):
            idx = re.findall(r'\[([^\[\]]*)\]', m.group(1))
            if len(idx) != self.rank:
                return False
            for (v, _), e in zip(self.vars, idx):
                if not re.fullmatch(r'\s*' + v + r'\s*(?:[+-]\s*[01]\s*)?', e):
                    return False
        # the array must not be reachable any other way (name used only as A[...])
        if re.search(r'\b' + name + r'\b(?!\s*\[)', self.body):
            return False
        return True

    def stencil_range(self, name):
        """Per loop axis, the (lowest, highest) offset at which the body reads
        `name` (each in -1..1).  Every index of [lo + lowest, last + highest]
        on an axis is read by some iteration, so a load clamped into that range
        stays inside the array."""
        lo = [None] * self.rank
        hi = [None] * self.rank
        for m in re.finditer(r'\b' + name + r'\s*((?:\[[^\[\]]*\]\s*)+)', self.body):
            idx = re.findall(r'\[([^\[\]]*)\]', m.group(1))
            for a, ((v, _), e) in enumerate(zip(self.vars, idx)):
                mm = re.fullmatch(r'\s*' + v + r'\s*(?:([+-])\s*([01])\s*)?', e)
                d = 0 if not mm.group(1) else (int(mm.group(2)) if mm.group(1) == '+' else -int(mm.group(2)))
                lo[a] = d if lo[a] is None else min(lo[a], d)
                hi[a] = d if hi[a] is None else max(hi[a], d)
        return [(l or 0, h or 0) for l, h in zip(lo, hi)]

    def tile_eligible(self):
        """Rank 3, every loop axis in the ACROSS clause with widths <= 1, and
        every access of the ACROSS array in the body of the form
        X[i + d0][j + d1][k + d2] with at most one offset, |offset| <= 1."""
        if not self.across or self.rank != 3:
            return False
        w = self.across['widths']
        if len(w) != 3 or any(lo > 1 or hi > 1 or (lo == 0 and hi == 0) for lo, hi in w):
            return False
        name = self.across['array']
        if re.search(r'\b' + name + r'\b(?!\s*\[)', self.body):
            return False
        for m in re.finditer(r'\b' + name + r'\s*((?:\[[^\[\]]*\]\s*)+)', self.body):
            idx = re.findall(r'\[([^\[\]]*)\]', m.group(1))
            if len(idx) != 3:
                return False
            moved = 0
            for (v, _), e in zip(self.vars, idx):
                mm = re.fullmatch(r'\s*' + v + r'\s*(?:([+-])\s*([01])\s*)?', e)
                if not mm:
                    return False
                if mm.group(2) == '1':
                    moved += 1
            if moved > 1:
                return False
        return True

    def chain_eligible(self):
        """Consecutive executions (sweeps) of the loop may run pipelined by a
        chain handler: a tiled-wavefront ACROSS loop marked for batching that
        writes no array but its ACROSS array and has scalar reductions only."""
        if not self.batch or not self.tile_eligible():
            return False
        X = self.across['array']
        for p in self.arrays:
            if p[1] != X and re.search(r'\b' + p[1] + r'\s*(?:\[[^\]]*\]\s*)+(?:=(?!=)|\+=|-=|\*=|/=|\+\+|--)', self.body):
                return False
        return all(self.rlen(red) == 1 for red in self.reds)

    def line_tile_eligible(self):
        """A line along the innermost axis can be staged in LDS and walked by
        one thread (the tiled line sweep)."""
        return self.line_stage_eligible(self.rank - 1)

    def line_stage_eligible(self, axis):
        """Rank 2 or 3, and every access of the ACROSS array in the body is
        X[i][j][k] with the loop's own variables, offset 0 on all axes but
        `axis`, |d| <= 1 there, and d = 0 where it is written: a thread's line
        along `axis` can then be staged (LDS or registers) chunk by chunk."""
        if not self.across or self.rank not in (2, 3):
            return False
        name = self.across['array']
        arr = [p for p in self.arrays if p[1] == name]
        if not arr or arr[0][3] != self.rank:
            return False
        if re.search(r'\b' + name + r'\b(?!\s*\[)', self.body):
            return False
        if re.search(r'(\+\+|--)\s*' + name + r'\b', self.body):
            return False
        for m in re.finditer(r'\b' + name + r'\s*((?:\[[^\[\]]*\]\s*)+)(\s*(?:=(?!=)|\+=|-=|\*=|/=|\+\+|--))?', self.body):
            idx = re.findall(r'\[([^\[\]]*)\]', m.group(1))
            if len(idx) != self.rank:
                return False
            for a, ((v, _), e) in enumerate(zip(self.vars, idx)):
                mm = re.fullmatch(r'\s*' + v + r'\s*(?:([+-])\s*([01])\s*)?', e)
                if not mm:
                    return False
                moved = mm.group(2) == '1'
                if moved and (a != axis or m.group(2)):
                    return False
        return True

    # static dependency masks the ACROSS clause can produce (identity correspondence)
    def dep_axes(self):
        if not self.across:
            return []
        return [a for a, (lo, hi) in enumerate(self.across['widths']) if lo > 0 or hi > 0]


# ----------------------------------------------------------------- host handlers
def seq_host_handler(L):
    """Sequential part of a region (converter_main.cpp:1310-1330, a loop of
    rank 0, host handler type 0): the statement runs once, by the calling
    thread, against the arrays' host copies."""
    formal = ['DvmType *pLoopRef']
    for p in L.params:
        formal.append(f'DvmType {p[1]}_hdr[]' if p[0] == 'array' else f'{p[2]} *{p[1]}_ptr')
    b = ['DvmType loop_ref = *pLoopRef;', 'DvmType device_num = dvmh_loop_get_device_num_C(loop_ref);']
    for p in L.params:
        if p[0] == 'array':
            _, nm, ct, ar = p
            dims = ''
            for j in range(2, ar + 1):
                dims += f'[{nm}_hdr[{j - 1}]/{nm}_hdr[{j}]]' if j < ar else f'[{nm}_hdr[{j - 1}]]'
            b.append(f'{ct} (*{nm}){dims} = ({ct} (*){dims})dvmh_get_natural_base_C(device_num, {nm}_hdr);')
        else:
            # the variable itself (a sequential part may assign the region's scalars)
            b.append(f'#define {p[1]} (*{p[1]}_ptr)')
    b.append('(void)loop_ref;')
    b.append('{\n' + indent(L.body + '\n', 1) + '}')
    for p in L.scalars:
        b.append(f'#undef {p[1]}')
    return f'void {L.name}_host({", ".join(formal)}) {{\n' + indent('\n'.join(b) + '\n', 1) + '}\n'


def host_handler(L):
    if L.sequential:
        return seq_host_handler(L)
    r = L.rank
    out = []
    formal = ['DvmType *pLoopRef']
    for p in L.params:
        formal.append(f'DvmType {p[1]}_hdr[]' if p[0] == 'array' else ('void *' if p[0] == 'carray' else f'{p[2]} *') + f'{p[1]}_ptr')
    out.append(f'void {L.name}_host({", ".join(formal)}) {{')
    b = []
    b.append('/* the handle this handler was called with, and the device it runs on */')
    b.append('DvmType loop_ref = *pLoopRef;')
    b.append('DvmType device_num = dvmh_loop_get_device_num_C(loop_ref);')
    b.append('/* array views and scalar addresses on that device */')
    for p in L.params:
        if p[0] == 'array':
            _, nm, ct, ar = p
            dims = ''
            for j in range(2, ar + 1):
                dims += f'[{nm}_hdr[{j - 1}]/{nm}_hdr[{j}]]' if j < ar else f'[{nm}_hdr[{j - 1}]]'
            b.append(f'{ct} (*{nm}){dims} = ({ct} (*){dims})dvmh_get_natural_base_C(device_num, {nm}_hdr);')
        elif p[0] == 'carray':
            # an ordinary C array of the program (loops outside regions run on the host)
            _, nm, ct, ext = p
            dims = ''.join(f'[{e}]' for e in ext[1:])
            b.append(f'{ct} (*{nm}){dims} = ({ct} (*){dims}){nm}_ptr;')
        else:
            b.append(f'{p[2]} {p[1]} = *{p[1]}_ptr;')
    b.append('/* portion bounds, host slot count, dependency bits */')
    b.append(f'DvmType boundsLow[{r}], boundsHigh[{r}], loopSteps[{r}];')
    b.append('int slotCount;')
    if L.across:
        b.append('DvmType dependencyMask;')
    b.append("/* the loop body's own variables: its indices, privates and reduction partials */")
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    for red in L.reds:
        n = L.rlen(red)
        b.append(f'{red[2]} {red[1]}{f"[{n}]" if n > 1 else ""};')
        if len(red) > 3:
            b.append(f'{red[4]} {red[3]};')
    b.append('')
    b.append('dvmh_loop_fill_bounds_C(loop_ref, boundsLow, boundsHigh, loopSteps);')
    b.append('slotCount = dvmh_loop_get_slot_count_C(loop_ref);')
    if L.across:
        b.append('dependencyMask = dvmh_loop_get_dependency_mask_C(loop_ref);')
        b.append('int slotReady[slotCount > 0 ? slotCount : 1];')
    privs = [v for v, _ in L.vars] + [re.split(r'[\s\*]+', pr.strip())[-1].split('[')[0] for pr in L.privates]
    redvars = []
    for red in L.reds:
        redvars.append(red[1])
        if len(red) > 3:
            redvars.append(red[3])
    clause = f'#pragma omp parallel num_threads(slotCount)'
    if privs:
        clause += f', private({", ".join(privs)})'
    if redvars:
        clause += f', private({", ".join(redvars)})'
    b.append(clause)
    b.append('{')
    inner = []
    if L.across:
        inner.append('int mySlot = 0, teamSize = slotCount;')
    for i, red in enumerate(L.reds):
        loc = ('&' + red[3]) if len(red) > 3 else '0'
        addr = red[1] if L.rlen(red) > 1 else '&' + red[1]
        inner.append(f'dvmh_loop_red_init_C(loop_ref, {i + 1}, {addr}, {loc});')

    def loop_head(i):
        v = L.vars[i][0]
        st = L.steps[i]
        cmp = '<=' if (st is None or not st.strip().startswith('-')) else '>='
        if st is None:
            cmp_expr = f'(loopSteps[{i}] > 0 ? {v} <= boundsHigh[{i}] : {v} >= boundsHigh[{i}])'
            return f'for ({v} = boundsLow[{i}]; {cmp_expr}; {v} += loopSteps[{i}])'
        inc = f'{v}++' if st == '1' else (f'{v}--' if st == '-1' else f'{v} += ({st})')
        return f'for ({v} = boundsLow[{i}]; {v} {cmp} boundsHigh[{i}]; {inc})'

    body = '{\n' + indent(L.body + '\n', 1) + '}'
    variants = L.rank + 1 if L.across else 1
    for par in range(variants):
        code = []
        if variants > 1:
            if par < L.rank:
                cond = f'if (((dependencyMask >> {L.rank - par - 1}) & 1) == 0) '
            else:
                cond = ''
            code.append(('} else ' if par > 0 else '') + cond + '{')
        depth = 1 if variants > 1 else 0
        lines = []
        if par < L.rank:
            for i in range(L.rank):
                if i == par:
                    # Under a sequential dependent axis every thread must get the
                    # same chunk each time (the "nowait" relies on it), so ACROSS
                    # loops use a static schedule; libgomp's runtime default is dynamic.
                    sched = 'static' if L.across else 'runtime'
                    lines.append(IND * i + f'#pragma omp for schedule({sched}), nowait')
                lines.append(IND * i + loop_head(i))
            lines.append(indent(body, L.rank))
        else:
            # every axis dependent: pipelined handshake over the first two axes
            lines.append('{')
            lines.append(IND + 'DvmType outerIters = (boundsHigh[1] - boundsLow[1]) / loopSteps[1] + 1;')
            lines.append(IND + 'if (outerIters < teamSize) teamSize = outerIters;')
            lines.append('#ifdef _OPENMP')
            lines.append(IND + 'teamSize = teamSize < omp_get_num_threads() ? teamSize : omp_get_num_threads();')
            lines.append(IND + 'mySlot = omp_get_thread_num();')
            lines.append('#endif')
            lines.append(IND + '__atomic_store_n(&slotReady[mySlot], 0, __ATOMIC_SEQ_CST);')
            lines.append(IND + '#pragma omp barrier')
            lines.append(IND + loop_head(0) + ' {')
            lines.append(IND * 2 + 'if (mySlot > 0 && mySlot < teamSize) {')
            lines.append(IND * 3 + 'while (!__atomic_load_n(&slotReady[mySlot - 1], __ATOMIC_ACQUIRE)) {}')
            lines.append(IND * 3 + '__atomic_store_n(&slotReady[mySlot - 1], 0, __ATOMIC_RELEASE);')
            lines.append(IND * 2 + '}')
            lines.append(IND * 2 + '#pragma omp for schedule(static), nowait')
            for i in range(1, L.rank):
                lines.append(IND * (i + 1) + loop_head(i))
            lines.append(indent(body, L.rank + 1))
            lines.append(IND * 2 + 'if (mySlot < teamSize - 1) {')
            lines.append(IND * 3 + 'while (__atomic_load_n(&slotReady[mySlot], __ATOMIC_ACQUIRE)) {}')
            lines.append(IND * 3 + '__atomic_store_n(&slotReady[mySlot], 1, __ATOMIC_RELEASE);')
            lines.append(IND * 2 + '}')
            lines.append(IND + '}')
            lines.append('}')
        code.append(indent('\n'.join(lines) + '\n', depth).rstrip('\n'))
        inner.extend(code)
    if variants > 1:
        inner.append('}')
    for i, red in enumerate(L.reds):
        loc = ('&' + red[3]) if len(red) > 3 else '0'
        addr = red[1] if L.rlen(red) > 1 else '&' + red[1]
        inner.append(f'dvmh_loop_red_post_C(loop_ref, {i + 1}, {addr}, {loc});')
    b.append(indent('\n'.join(inner) + '\n', 1).rstrip('\n'))
    b.append('}')
    out.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    out.append('}')
    return '\n'.join(out) + '\n'


# ----------------------------------------------------------------- GPU: common pieces
def scalar_bind(body, p, pre='', ptr=None):
    """A scalar parameter inside a kernel: a const copy when the body never
    assigns it or takes its address (loads and what is computed from it are
    then loop-invariant to the compiler), else a reference to the device copy."""
    nm = p[1]
    written = re.search(r'(?<![\w.>])' + nm + r'\s*(?:=(?!=)|\+=|-=|\*=|/=|%=|&=|\|=|\^=|<<=|>>=|\+\+|--)'
                        r'|(?:\+\+|--)\s*' + nm + r'\b|&\s*' + nm + r'\b', body)
    src = ptr if ptr else f'{pre}{nm}_ptr'
    if written:
        return f'{p[2]} &{nm} = *{src};'
    return f'const {p[2]} {nm} = *{src};'


def kernel_array_params(L, unit):
    fp, setup = [], []
    for p in L.arrays:
        _, nm, ct, ar = p
        fp.append(f'{ct} *__restrict__ {nm}_base')
        fp += [f'IndexT {nm}_s{k}' for k in range(ar)]
        strides = ', '.join(f'{nm}_s{k}' for k in range(ar))
        setup.append(f'const IndexT {nm}_str[{ar}] = {{{strides}}};')
        setup.append(f'const auto {nm} = dvmh_dev::makeView<{ct}, IndexT, {ar}, {"true" if unit else "false"}>({nm}_base, {nm}_str);')
    for p in L.scalars:
        fp.append(f'{p[2]} *__restrict__ {p[1]}_ptr')
        setup.append(scalar_bind(L.body, p))
    return fp, setup


def kernel_diag_array_params(L):
    fp, setup = [], []
    for p in L.arrays:
        _, nm, ct, ar = p
        fp.append(f'{ct} *__restrict__ {nm}_base')
        fp += [f'IndexT {nm}_s{k}' for k in range(ar)]
        fp.append(f'dvmh_dev::DiagInfo<IndexT> {nm}_di')
        strides = ', '.join(f'{nm}_s{k}' for k in range(ar))
        setup.append(f'const IndexT {nm}_str[{ar}] = {{{strides}}};')
        setup.append(f'const auto {nm} = dvmh_dev::makeDiagView<{ct}, IndexT, {ar}>({nm}_base, {nm}_str, {nm}_di);')
    for p in L.scalars:
        fp.append(f'{p[2]} *__restrict__ {p[1]}_ptr')
        setup.append(scalar_bind(L.body, p))
    return fp, setup


def red_kernel_params(L):
    fp = []
    for red in L.reds:
        n = L.rlen(red)
        if n > 1:
            # a reduction array: its initial values by value; element q's
            # partial of workgroup slot s at grid[q * items + s]
            fp += [f'dvmh_dev::RedVec<{red[2]}, {n}> {red[1]}_init', f'{red[2]} *{red[1]}_grid', f'long {red[1]}_items']
            continue
        fp += [f'{red[2]} {red[1]}_init', f'{red[2]} *{red[1]}_grid']
        if len(red) > 3:
            fp += [f'{red[4]} {red[3]}_init', f'{red[4]} *{red[3]}_grid']
    return fp


def red_kernel_locals(L):
    s = []
    for red in L.reds:
        n = L.rlen(red)
        if n > 1:
            s.append(f'{red[2]} {red[1]}[{n}];')
            s.append(f'for (int dvmh_q = 0; dvmh_q < {n}; dvmh_q++) {red[1]}[dvmh_q] = {red[1]}_init.v[dvmh_q];')
            continue
        s.append(f'{red[2]} {red[1]} = {red[1]}_init;')
        if len(red) > 3:
            s.append(f'{red[4]} {red[3]} = {red[3]}_init;')
    return s


def red_kernel_tail(L, slot, accumulate):
    """Workgroup reduction; thread 0 stores (or folds into) the workgroup's slot."""
    s = []
    tid0 = '(threadIdx.x | threadIdx.y | threadIdx.z) == 0'
    for red in L.reds:
        f, v, ct = red[0], red[1], red[2]
        n = L.rlen(red)
        if n > 1:
            s.append(f'for (int dvmh_q = 0; dvmh_q < {n}; dvmh_q++) {{')
            s.append(f'    {ct} dvmh_r = dvmh_dev::blockReduce({v}[dvmh_q], dvmh_dev::{RED_OP[f]}());')
            cell = f'{v}_grid[(long)dvmh_q * {v}_items + ({slot})]'
            if accumulate:
                s.append(f'    if ({tid0}) {cell} = dvmh_dev::{RED_OP[f]}()({cell}, dvmh_r);')
            else:
                s.append(f'    if ({tid0}) {cell} = dvmh_r;')
            s.append('}')
            continue
        if len(red) > 3:
            loc = red[3]
            mx = 'true' if f == 'maxloc' else 'false'
            s.append(f'dvmh_dev::blockReduceLoc<{ct}, {red[4]}, {mx}>({v}, {loc});')
            if accumulate:
                cmpo = '<' if f == 'maxloc' else '>'
                s.append(f'if ({tid0}) {{')
                s.append(f'    {ct} ov = {v}_grid[{slot}];')
                s.append(f'    {red[4]} ol = {loc}_grid[{slot}];')
                s.append(f'    if (ov {cmpo} {v} || (!({v} {cmpo} ov) && {loc} < ol)) {{ {v}_grid[{slot}] = {v}; {loc}_grid[{slot}] = {loc}; }}')
                s.append('}')
            else:
                s.append(f'if ({tid0}) {{ {v}_grid[{slot}] = {v}; {loc}_grid[{slot}] = {loc}; }}')
        else:
            s.append(f'{v} = dvmh_dev::blockReduce({v}, dvmh_dev::{RED_OP[f]}());')
            if accumulate:
                s.append(f'if ({tid0}) {v}_grid[{slot}] = dvmh_dev::{RED_OP[f]}()({v}_grid[{slot}], {v});')
            else:
                s.append(f'if ({tid0}) {v}_grid[{slot}] = {v};')
    return s


def regs_query(kernel, loop_ref='loop_ref'):
    """Handler lines that ask once for the registers per thread of its main
    kernel and hand them to dvmh_loop_cuda_get_config_C."""
    return ['static int dvmh_regs = -1;   /* registers per thread of the main kernel, asked once */',
            f'if (dvmh_regs < 0) dvmh_regs = dvmh_dev::kernelRegs(reinterpret_cast<const void *>(&{kernel}));',
            f'dvmh_loop_cuda_get_config_C({loop_ref}, 0, dvmh_regs, &threads, &stream, 0);']


def handler_prologue(L, kind_const, want_ext=False, autotfm=False, regs_kernel=None):
    h = []
    h.append('DvmType tmpVar;')
    h.append('/* the handle this handler was called with, and the device it runs on */')
    h.append('DvmType loop_ref = *pLoopRef;')
    h.append('DvmType device_num = dvmh_loop_get_device_num_C(loop_ref);')
    h.append('DvmType tune[DVMH_AMD_TUNE_N];')
    h.append(f'dvmh_amd_loop_get_tuning_C(loop_ref, {kind_const}, tune);')
    h.append('/* array views and scalar addresses on that device */')
    for p in L.arrays:
        _, nm, ct, ar = p
        if autotfm:
            h.append(f'if (tune[DVMH_AMD_T_TRANSFORM]) dvmh_loop_autotransform_C(loop_ref, {nm}_hdr);')
        h.append(f'{ct} *{nm} = ({ct} *)dvmh_get_natural_base_C(device_num, {nm}_hdr);')
        h.append(f'DvmType {nm}_devHdr[64 + 7];')
        ext = f'{nm}_devHdr + 64' if want_ext else '0'
        h.append(f'tmpVar = dvmh_fill_header_C(device_num, {nm}, {nm}_hdr, {nm}_devHdr, {ext});')
        h.append(f'int {nm}_tfm = (int)tmpVar;')
    for p in L.scalars:
        h.append(f'{p[2]} *{p[1]} = ({p[2]} *)dvmh_get_device_addr_C(device_num, {p[1]}_ptr);')
    h.append(f'DvmType boundsLow[{L.rank}], boundsHigh[{L.rank}], loopSteps[{L.rank}];')
    h.append('dim3 threads(0, 0, 0);')
    h.append('hipStream_t stream;')
    h.append('/* 32- or 64-bit index arithmetic, from the largest extent involved */')
    h.append('int kernelIndexT = (int)dvmh_loop_guess_index_type_C(loop_ref);')
    h.append('if (kernelIndexT == rt_LONG) kernelIndexT = rt_LLONG;')
    h += regs_query(regs_kernel)
    h.append('dvmh_loop_fill_bounds_C(loop_ref, boundsLow, boundsHigh, loopSteps);')
    return h


def handler_reds_prepare(L, count_expr, fill):
    h = []
    for i, red in enumerate(L.reds):
        n = L.rlen(red)
        if n > 1:
            h.append(f'{red[2]} {red[1]}[{n}];')
            h.append(f'{red[2]} *{red[1]}_grid;')
            h.append(f'long {red[1]}_items = (long)({count_expr});')
            h.append(f'dvmh_loop_cuda_register_red_C(loop_ref, {i + 1}, (void **)&{red[1]}_grid, 0);')
            h.append(f'dvmh_loop_red_init_C(loop_ref, {i + 1}, {red[1]}, 0);')
            h.append(f'dvmh_loop_cuda_red_prepare_C(loop_ref, {i + 1}, {count_expr}, {fill});')
            continue
        h.append(f'{red[2]} {red[1]};')
        h.append(f'{red[2]} *{red[1]}_grid;')
        if len(red) > 3:
            h.append(f'{red[4]} {red[3]};')
            h.append(f'{red[4]} *{red[3]}_grid;')
            h.append(f'dvmh_loop_cuda_register_red_C(loop_ref, {i + 1}, (void **)&{red[1]}_grid, (void **)&{red[3]}_grid);')
            h.append(f'dvmh_loop_red_init_C(loop_ref, {i + 1}, &{red[1]}, &{red[3]});')
        else:
            h.append(f'dvmh_loop_cuda_register_red_C(loop_ref, {i + 1}, (void **)&{red[1]}_grid, 0);')
            h.append(f'dvmh_loop_red_init_C(loop_ref, {i + 1}, &{red[1]}, 0);')
        h.append(f'dvmh_loop_cuda_red_prepare_C(loop_ref, {i + 1}, {count_expr}, {fill});')
    return h


def handler_reds_finish(L):
    return [f'dvmh_loop_cuda_red_finish_C(loop_ref, {i + 1});' for i in range(len(L.reds))]


def red_fact_args(L):
    a = []
    for red in L.reds:
        n = L.rlen(red)
        if n > 1:
            a += [f'dvmh_dev::redVec<{red[2]}, {n}>({red[1]})', f'{red[1]}_grid', f'{red[1]}_items']
            continue
        a += [red[1], f'{red[1]}_grid']
        if len(red) > 3:
            a += [red[3], f'{red[3]}_grid']
    return a


def array_fact_args(L, itype, diag=False):
    a = []
    for p in L.arrays:
        _, nm, ct, ar = p
        a.append(nm)
        a += [f'({itype}){nm}_devHdr[{k + 1}]' for k in range(ar)]
        if diag:
            a.append(f'{nm}_di_{itype.replace(" ", "_")}')
    for p in L.scalars:
        a.append(p[1])
    return a


def var_assign(L, i, tname, lo, st):
    v, ct = L.vars[i]
    return f'{v} = ({ct})({lo} + {tname} * {st});'


def launch_dispatch(kernels, args, grid, block='threads'):
    """kernels: list of (condition, kernel instantiation text)."""
    s = []
    first = True
    for cond, k in kernels:
        kw = 'if' if first else 'else if'
        first = False
        s.append(f'{kw} ({cond}) hipLaunchKernelGGL(({k}), {grid}, {block}, 0, stream, {", ".join(args)});')
    return s


# ----------------------------------------------------------------- GPU: PARALLEL
def parallel_kernel(L):
    r = L.rank
    fp, setup = kernel_array_params(L, unit=True)
    for i in range(r):
        fp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    fp += ['IndexT nbx', 'IndexT nby', 'IndexT nbz', 'long nblocks', 'int xcd']
    fp += red_kernel_params(L)
    k = []
    k.append(f'template <typename IndexT, int DEPTH>')
    k.append(f'__global__ __launch_bounds__(1024) void {L.name}_kernel({", ".join(fp)}) {{')
    b = list(setup)
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    b += red_kernel_locals(L)
    b.append('long tile = dvmh_dev::xcdTile((long)blockIdx.x, nblocks, xcd);')
    b.append('const IndexT bx = (IndexT)(tile % nbx); tile /= nbx;')
    b.append('const IndexT by = (IndexT)(tile % nby); tile /= nby;')
    b.append('const IndexT bz = (IndexT)(tile % nbz); tile /= nbz;')
    # thread-mapped axes
    ax_x = r - 1
    ax_y = r - 2 if r >= 2 else None
    ax_z = r - 3 if r >= 3 else None
    dep_ax = ax_z if ax_z is not None else (ax_y if ax_y is not None else ax_x)
    outer = list(range(0, r - 3)) if r > 3 else []
    for a in reversed(outer):
        b.append(f'const IndexT t{a} = (IndexT)(tile % cnt{a}); tile /= cnt{a};')
    b.append(f'const IndexT t{ax_x}_0 = bx * (IndexT)blockDim.x + (IndexT)threadIdx.x;')
    if ax_y is not None:
        b.append(f'const IndexT t{ax_y}_0 = by * (IndexT)blockDim.y + (IndexT)threadIdx.y;')
    if ax_z is not None:
        b.append(f'const IndexT t{ax_z}_0 = bz * (IndexT)blockDim.z + (IndexT)threadIdx.z;')
    conds = []
    for a in [ax_x, ax_y, ax_z]:
        if a is None:
            continue
        if a == dep_ax and r >= 2:
            b.append(f'const IndexT t{a}_b = t{a}_0 * DEPTH;')
        else:
            b.append(f'const IndexT t{a} = t{a}_0;')
            conds.append(f't{a} < cnt{a}')
    if r >= 2:
        # Threads whose DEPTH iterations are all inside the loop take an
        # unguarded copy: one basic block, so the loads of every iteration can
        # be issued before the first result is stored (iterations of a
        # PARALLEL loop are independent and array pointers are restrict).
        # Threads on the edges take the guarded copy.
        whole = ' && '.join([f't{dep_ax}_b + DEPTH <= cnt{dep_ax}'] + conds)
        b.append(f'if ({whole}) {{')
        fast = [f'const IndexT t{dep_ax} = t{dep_ax}_b + dd;']
        for i in range(r):
            fast.append(var_assign(L, i, f't{i}', f'lo{i}', f'st{i}'))
        fast.append('{\n' + indent(L.body + '\n', 1) + '}')
        b.append(IND + '#pragma unroll')
        b.append(IND + 'for (int dd = 0; dd < DEPTH; dd++) {')
        b += [IND * 2 + ln for ln in '\n'.join(fast).split('\n')]
        b.append(IND + '}')
        b.append('} else {')
        b.append(IND + '#pragma unroll')
        b.append(IND + 'for (int dd = 0; dd < DEPTH; dd++) {')
        inner = [f'const IndexT t{dep_ax} = t{dep_ax}_b + dd;']
        conds2 = conds + [f't{dep_ax} < cnt{dep_ax}']
    else:
        inner = []
        conds2 = conds
    inner.append(f'if ({" && ".join(conds2)}) {{')
    for i in range(r):
        inner.append(IND + var_assign(L, i, f't{i}', f'lo{i}', f'st{i}'))
    inner.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 1))
    inner.append('}')
    if r >= 2:
        b += [IND * 2 + ln for ln in '\n'.join(inner).split('\n')]
        b.append(IND + '}')
        b.append('}')
    else:
        b += inner
    b += red_kernel_tail(L, 'blockIdx.x', accumulate=False)
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    return '\n'.join(k) + '\n'


# Register slots per thread for plane loads (NPF) instantiated for each COLS:
# a launch needs ceil(plane / threads) slots, between COLS + 1 (rank 2, or a
# tall tile) and 3 * COLS + 2 (rank 3, one row of threads), and takes the
# smallest instantiated value that covers it.
MARCH_NPF = {1: [2, 3, 4, 5], 2: [3, 4, 5, 7, 8], 3: [4, 5, 7, 10, 11], 4: [5, 7, 10, 14]}


def march_kernel(L, tail=None):
    """LDS march variant of a PARALLEL loop with stencil arrays (rank 2 or 3,
    unit steps).  A workgroup owns a tile of the inner axes -- blockDim.x * COLS
    points of the innermost axis (x), blockDim.y of the middle one for rank 3 --
    and marches `march` planes along the outermost axis.  Each stencil array is
    staged in LDS as a rolling window of three planes (tile plus a halo of one),
    every element loaded from HBM once per tile; the body reads it through a
    MarchView with the same A[i][j][k] syntax, other arrays stay in HBM.

    With `tail` (a loop fused after L, see pair_eligible), only the tile
    function is returned, named after the pair: after L's body, each point not
    on an edge of its tile that faces another tile runs the tail loop's body at
    the same point (pair_tail)."""
    r = L.rank
    ax_x, ax_y = r - 1, (r - 2 if r == 3 else None)
    fp, setup = [], []
    for p in L.arrays:
        _, nm, ct, ar = p
        fp.append(f'{ct} *__restrict__ {nm}_base')
        fp += [f'IndexT {nm}_s{k}' for k in range(ar)]
        strides = ', '.join(f'{nm}_s{k}' for k in range(ar))
        if nm in L.stencil:
            continue   # read through LDS; plane loads use the base and strides directly
        setup.append(f'const IndexT {nm}_str[{ar}] = {{{strides}}};')
        setup.append(f'const auto {nm} = dvmh_dev::makeView<{ct}, IndexT, {ar}, true>({nm}_base, {nm}_str);')
    for p in L.scalars:
        fp.append(f'{p[2]} *__restrict__ {p[1]}_ptr')
        setup.append(scalar_bind(L.body, p))
    for i in range(r):
        fp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    fp += ['IndexT nbx', 'IndexT nby', 'IndexT nbz', 'IndexT march']
    # the march over one tile is a device function: the per-loop kernel calls it
    # for its workgroup's tile, the batch kernel for every tile it walks
    # xshift: the x-tiling starts xshift points before the loop's first point
    # (tile 0 is that much narrower), so that later tiles begin where the
    # written array's rows are 128-byte aligned; batch and pair tiles pass 0
    tp = list(fp) + ['long tile', 'IndexT xshift']
    for red in L.reds:
        n = L.rlen(red)
        tp.append(f'{red[2]} (&{red[1]})[{n}]' if n > 1 else f'{red[2]} &{red[1]}')
        if len(red) > 3:
            tp.append(f'{red[4]} &{red[3]}')
    fname = f'{L.name}_march_tile'
    if tail is not None:
        tp += pair_tail_params(L, tail)
        fname = f'{pair_name(L, tail)}_pair_tile'
    k = ['template <typename IndexT, int COLS, int NPF>',
         f'__device__ static inline void {fname}({", ".join(tp)}) {{']
    b = ['extern __shared__ __align__(16) unsigned char dvmh_lds[];']
    b += setup
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    b.append('const IndexT bx = (IndexT)(tile % nbx); tile /= nbx;')
    b.append('const IndexT by = (IndexT)(tile % nby); tile /= nby;')
    b.append('const IndexT bz = (IndexT)tile;')
    b.append('const IndexT tileX = (IndexT)blockDim.x * COLS;')
    b.append(f'const IndexT tx0 = bx * tileX - xshift;')
    if ax_y is not None:
        b.append('const IndexT tileY = (IndexT)blockDim.y;')
        b.append('const IndexT ty0 = by * tileY;')
        b.append('const IndexT rows = tileY + 2;')
    else:
        b.append('const IndexT rows = 1;')
    b.append('const IndexT cols = tileX + 2;')
    b.append('const IndexT plane = rows * cols;')
    # LDS: three planes per stencil array, 16-byte aligned
    off = '0'
    for nm in L.stencil:
        ct = [p for p in L.arrays if p[1] == nm][0][2]
        b.append(f'{ct} *{nm}_lds = ({ct} *)(dvmh_lds + {off});')
        off = f'((({off}) + 3 * plane * (IndexT)sizeof({ct}) + 15) & ~(IndexT)15)'
    b.append(f'const IndexT kx0 = lo{ax_x} + tx0 - 1;   // array index held by LDS column 0')
    if ax_y is not None:
        b.append(f'const IndexT jy0 = lo{ax_y} + ty0 - 1;   // array index held by LDS row 0')
    b.append('const IndexT tz0 = bz * march;')
    b.append('const IndexT nMarch = cnt0 - tz0 < march ? cnt0 - tz0 : march;')
    b.append('const IndexT i0 = lo0 + tz0;')
    b.append('const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);')
    b.append('const int nthr = (int)(blockDim.x * blockDim.y);')
    # Plane loads are software-pipelined through registers: each thread owns
    # plane elements e = tid + m * nthr (m < NPF; the handler instantiates NPF
    # just above ceil(plane / threads), so no registers are held for elements
    # a shape never has); issue() loads a plane from HBM into
    # registers, commit() stores them into an LDS slot.  The next plane's loads
    # are issued before the current plane is computed, so their latency is
    # hidden behind the compute.  Only indices some point of the loop reads
    # (the loop box plus the halo) are loaded.  Which elements a thread owns,
    # their offsets inside a plane and whether they are loaded do not change
    # along the march: they are worked out once, and a plane load is a base
    # pointer plus those offsets.
    b.append('unsigned inPlane = 0;   // bit m: element tid + m * nthr lies inside the plane')
    for nm in L.stencil:
        rng = L.stencil_range(nm)
        b.append(f'// {nm}: indices the body reads, per axis (a load is clamped into them)')
        for a in range(r):
            dl, dh = rng[a]
            b.append(f'const IndexT {nm}_lo{a} = lo{a} + ({dl}), {nm}_hi{a} = lo{a} + cnt{a} - 1 + ({dh});')
        b.append(f'IndexT {nm}_off[NPF];')
    lines = ['{',
             IND + 'IndexT rr = (IndexT)tid / cols, cc = (IndexT)tid - rr * cols;',
             IND + 'const IndexT dR = (IndexT)nthr / cols, dC = (IndexT)nthr - dR * cols;',
             IND + '#pragma unroll',
             IND + 'for (int m = 0; m < NPF; m++) {',
             IND * 2 + 'const IndexT kk = kx0 + cc;',
             IND * 2 + 'inPlane |= ((IndexT)tid + (IndexT)m * nthr < plane ? 1u : 0u) << m;']
    if ax_y is not None:
        lines.append(IND * 2 + 'const IndexT jj = jy0 + rr;')
    for nm in L.stencil:
        lines.append(IND * 2 + f'const IndexT {nm}_k = kk < {nm}_lo{ax_x} ? {nm}_lo{ax_x} : (kk > {nm}_hi{ax_x} ? {nm}_hi{ax_x} : kk);')
        if ax_y is not None:
            lines.append(IND * 2 + f'const IndexT {nm}_j = jj < {nm}_lo{ax_y} ? {nm}_lo{ax_y} : (jj > {nm}_hi{ax_y} ? {nm}_hi{ax_y} : jj);')
            lines.append(IND * 2 + f'{nm}_off[m] = {nm}_j * {nm}_s1 + {nm}_k;')
        else:
            lines.append(IND * 2 + f'{nm}_off[m] = {nm}_k;')
    lines += [IND * 2 + 'cc += dC;',
              IND * 2 + 'rr += dR;',
              IND * 2 + 'if (cc >= cols) { cc -= cols; rr++; }',
              IND + '}',
              '}']
    for nm in L.stencil:
        ct = [p for p in L.arrays if p[1] == nm][0][2]
        lines.append(f'{ct} {nm}_pf[NPF];')
    # Elements outside the indices the body reads are loaded from a clamped
    # (valid) address; no iteration reads their LDS copies.
    lines += ['auto issue = [&](IndexT ip) {']
    for nm in L.stencil:
        ct = [p for p in L.arrays if p[1] == nm][0][2]
        lines.append(IND + f'const IndexT {nm}_i = ip < {nm}_lo0 ? {nm}_lo0 : (ip > {nm}_hi0 ? {nm}_hi0 : ip);')
        lines.append(IND + f'const {ct} *{nm}_pl = {nm}_base + {nm}_i * {nm}_s0;')
    lines += [IND + '#pragma unroll',
              IND + 'for (int m = 0; m < NPF; m++) {']
    lines.append(IND * 2 + 'const bool ld = (inPlane >> m) & 1u;')
    for nm in L.stencil:
        ct = [p for p in L.arrays if p[1] == nm][0][2]
        lines.append(IND * 2 + f'{nm}_pf[m] = ld ? {nm}_pl[{nm}_off[m]] : ({ct})0;')
    lines += [IND + '}',
              '};',
              'auto commit = [&](int slot) {']
    for nm in L.stencil:
        ct = [p for p in L.arrays if p[1] == nm][0][2]
        lines.append(IND + f'{ct} *{nm}_dst = {nm}_lds + (IndexT)slot * plane + tid;')
    lines += [IND + '#pragma unroll',
              IND + 'for (int m = 0; m < NPF; m++) {',
              IND * 2 + 'if ((inPlane >> m) & 1u) {']
    for nm in L.stencil:
        lines.append(IND * 3 + f'{nm}_dst[(IndexT)m * nthr] = {nm}_pf[m];')
    lines += [IND * 2 + '}', IND + '}', '};']
    b += lines
    b.append('issue(i0 - 1);')
    b.append('commit(0);')
    b.append('issue(i0);')
    b.append('commit(1);')
    b.append('issue(i0 + 1);')
    b.append('for (IndexT d = 0; d < nMarch; d++) {')
    inner = ['commit((int)((d + 2) % 3));   // plane i0 + d + 1',
             '__syncthreads();',
             'if (d + 1 < nMarch) issue(i0 + d + 2);   // in flight while plane d is computed',
             'const int sCur = (int)((d + 1) % 3);']
    for nm in L.stencil:
        ct = [p for p in L.arrays if p[1] == nm][0][2]
        if ax_y is not None:
            inner.append(f'const dvmh_dev::MarchView3<{ct}, IndexT> {nm}{{{nm}_lds, rows, cols, jy0, kx0, i0 + d, sCur}};')
        else:
            inner.append(f'const dvmh_dev::MarchView2<{ct}, IndexT> {nm}{{{nm}_lds, cols, kx0, i0 + d, sCur}};')
    inner.append('const IndexT t0 = tz0 + d;')
    if ax_y is not None:
        inner.append(f'const IndexT t{ax_y} = ty0 + (IndexT)threadIdx.y;')
    inner.append('#pragma unroll')
    inner.append('for (int c = 0; c < COLS; c++) {')
    pt = [f'const IndexT t{ax_x} = tx0 + (IndexT)threadIdx.x + (IndexT)c * (IndexT)blockDim.x;']
    cc = [f't{ax_x} >= 0', f't{ax_x} < cnt{ax_x}'] + ([f't{ax_y} < cnt{ax_y}'] if ax_y is not None else [])
    pt.append(f'if ({" && ".join(cc)}) {{')
    for i in range(r):
        pt.append(IND + var_assign(L, i, f't{i}', f'lo{i}', '1'))   # the variant runs unit steps only
    pt.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 1))
    if tail is not None:
        pt += [IND + ln for ln in pair_tail(L, tail, ax_x, ax_y)]
    pt.append('}')
    inner += [IND + ln for ln in '\n'.join(pt).split('\n')]
    inner.append('}')
    inner.append('__syncthreads();')
    b += [IND + ln for ln in '\n'.join(inner).split('\n')]
    b.append('}')
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    if tail is not None:
        return '\n'.join(k) + '\n'
    # the per-loop kernel: one tile per workgroup
    names = [p_[1] for p_ in L.params]
    kp = list(fp[:-4]) + ['IndexT nbx', 'IndexT nby', 'IndexT nbz', 'long nblocks', 'int xcd', 'IndexT march', 'IndexT xshift'] + red_kernel_params(L)
    call = []
    for p_ in L.arrays:
        _, nm, ct, ar = p_
        call += [f'{nm}_base'] + [f'{nm}_s{q}' for q in range(ar)]
    for p_ in L.scalars:
        call.append(f'{p_[1]}_ptr')
    for i in range(r):
        call += [f'lo{i}', f'cnt{i}', f'st{i}']
    call += ['nbx', 'nby', 'nbz', 'march', 'dvmh_dev::xcdTile((long)blockIdx.x, nblocks, xcd)', 'xshift']
    for red in L.reds:
        call.append(red[1])
        if len(red) > 3:
            call.append(red[3])
    g = ['template <typename IndexT, int COLS, int NPF>',
         f'__global__ __launch_bounds__(1024) void {L.name}_march_kernel({", ".join(kp)}) {{']
    gb = red_kernel_locals(L)
    gb.append(f'{L.name}_march_tile<IndexT, COLS, NPF>({", ".join(call)});')
    gb += red_kernel_tail(L, 'blockIdx.x', accumulate=False)
    g.append(indent('\n'.join(gb) + '\n', 1).rstrip('\n'))
    g.append('}')
    return '\n'.join(k) + '\n\n' + '\n'.join(g) + '\n'


def parallel_handler(L):
    r = L.rank
    h = [f'extern "C" void {L.name}_cuda(DvmType *pLoopRef{"".join(", DvmType " + p[1] + "_hdr[]" if p[0] == "array" else ", " + p[2] + " *" + p[1] + "_ptr" for p in L.params)}) {{']
    b = handler_prologue(L, 'DVMH_AMD_KIND_STENCIL' if L.stencil else 'DVMH_AMD_KIND_PARALLEL',
                         regs_kernel=f'{L.name}_kernel<int, 1>')
    for p in L.arrays:
        b.append(f'if ({p[1]}_tfm != 0) {{ fprintf(stderr, "{L.name}: array {p[1]} is not in natural layout\\n"); abort(); }}')
    b.append('int depth = (int)tune[DVMH_AMD_T_DEPTH];')
    b.append('if (depth != 1 && depth != 2 && depth != 4 && depth != 8) depth = 1;')
    for i in range(r):
        b.append(f'DvmType cnt{i} = (boundsHigh[{i}] - boundsLow[{i}]) / loopSteps[{i}] + 1;')
    ax_x, ax_y, ax_z = r - 1, (r - 2 if r >= 2 else None), (r - 3 if r >= 3 else None)
    dep_ax = ax_z if ax_z is not None else (ax_y if ax_y is not None else ax_x)
    if L.stencil:
        # LDS march variant: COLS points per thread on x, `march` planes per workgroup
        b.append('int variant = (int)tune[DVMH_AMD_T_VARIANT];')
        b.append('int mcols = (int)tune[DVMH_AMD_T_TILE_A];')
        b.append('DvmType march = tune[DVMH_AMD_T_CHUNK];')
        b.append('size_t ldsBytes = 0;')
        b.append('int npf = 0;')
        unit = ' && '.join(f'loopSteps[{i}] == 1' for i in range(r))
        shape_ok = 'threads.z == 1' + (' && threads.y == 1' if r == 2 else '')
        b.append(f'if (variant == 1 && ({unit}) && {shape_ok} && mcols >= 1 && mcols <= 4 && march >= 1) {{')
        rows = '((size_t)threads.y + 2)' if r == 3 else '(size_t)1'
        b.append(IND + f'size_t mplane = {rows} * ((size_t)threads.x * mcols + 2);')
        for nm in L.stencil:
            ct = [p for p in L.arrays if p[1] == nm][0][2]
            b.append(IND + f'ldsBytes += (3 * mplane * sizeof({ct}) + 15) & ~(size_t)15;')
        b.append(IND + 'if (ldsBytes > 64512) variant = 0;   // leave room for the reductions\' static LDS')
        b.append(IND + 'size_t mthreads = (size_t)threads.x * threads.y;')
        b.append(IND + '// register share of a plane: the smallest instantiated NPF that covers it')
        b.append(IND + 'size_t need = (mplane + mthreads - 1) / mthreads;')
        for cols, sets in MARCH_NPF.items():
            b.append(IND + f'if (mcols == {cols}) {{ ' + ' '.join(f'if (npf == 0 && need <= {v}) npf = {v};' for v in sets) + ' }')
        b.append(IND + 'if (npf == 0 || kernelIndexT != rt_INT) variant = 0;   // 64-bit indices: the direct kernel')
        b.append('} else {')
        b.append(IND + 'variant = 0;')
        b.append('}')
        b.append('DvmType nbx, nby, nbz;')
        b.append('DvmType xshift = 0;')
        b.append('if (variant == 1) {')
        # tiles after the first start where the rows of the array the body
        # writes are 128-byte aligned (full cache lines stored)
        wr = [p for p in L.arrays if p[1] not in L.stencil and p[3] == r and
              re.search(r'\b' + p[1] + r'\s*(?:\[[^\]]*\]\s*)+(?:=(?!=)|\+=|-=|\*=|/=)', L.body)]
        if wr:
            _, wn, wct, _ = wr[0]
            addr = ' + '.join([wn] + [f'boundsLow[{k}] * {wn}_devHdr[{k + 1}]' for k in range(r)])
            b.append(IND + f'{{ const uintptr_t a = (uintptr_t)({addr});   /* the first point of {wn} */')
            b.append(IND + f'  if ({wn}_devHdr[{r}] == 1 && a % sizeof({wct}) == 0) xshift = (DvmType)((a / sizeof({wct})) % (128 / sizeof({wct}))); }}')
        b.append(IND + f'nbx = (cnt{ax_x} + xshift + (DvmType)threads.x * mcols - 1) / ((DvmType)threads.x * mcols);')
        b.append(IND + (f'nby = (cnt{ax_y} + threads.y - 1) / threads.y;' if r == 3 else 'nby = 1;'))
        b.append(IND + 'nbz = (cnt0 + march - 1) / march;')
        b.append('} else {')
        pre = IND
    else:
        b.append('DvmType nbx, nby, nbz;')
        pre = ''
    b.append(pre + f'nbx = (cnt{ax_x} + threads.x - 1) / threads.x;')
    if ax_y is not None:
        dy = ' * depth' if dep_ax == ax_y else ''
        b.append(pre + f'nby = (cnt{ax_y} + threads.y{dy} - 1) / (threads.y{dy});')
    else:
        b.append(pre + 'nby = 1;')
    if ax_z is not None:
        b.append(pre + f'nbz = (cnt{ax_z} + threads.z * depth - 1) / (threads.z * depth);')
    else:
        b.append(pre + 'nbz = 1;')
    if L.stencil:
        b.append('}')
    nb = 'nbx * nby * nbz' + ''.join(f' * cnt{a}' for a in range(0, r - 3))
    b.append(f'DvmType nblocks = {nb};')
    b += handler_reds_prepare(L, 'nblocks', 0)
    kernels = []
    for it in ['int', 'long long']:
        for dp in [1, 2, 4, 8]:
            cond = f'kernelIndexT == {"rt_INT" if it == "int" else "rt_LLONG"} && depth == {dp}'
            kernels.append((cond, f'{L.name}_kernel<{it}, {dp}>'))
    b.append('if (nblocks > 0) {')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 0);')
    if L.stencil:
        b.append(IND + 'if (variant == 1) {')
        it = 'int'
        args = array_fact_args(L, it)
        for i in range(r):
            args += [f'({it})boundsLow[{i}]', f'({it})cnt{i}', f'({it})loopSteps[{i}]']
        args += [f'({it})nbx', f'({it})nby', f'({it})nbz', 'nblocks', '(int)tune[DVMH_AMD_T_XCD]', f'({it})march', f'({it})xshift']
        args += red_fact_args(L)
        for cols, sets in MARCH_NPF.items():
            for v in sets:
                b.append(IND * 2 + f'if (mcols == {cols} && npf == {v}) hipLaunchKernelGGL(({L.name}_march_kernel<{it}, {cols}, {v}>), '
                         f'dim3((unsigned)nblocks), threads, ldsBytes, stream, {", ".join(args)});')
        b.append(IND + '} else {')
        ind2 = IND * 2
    else:
        ind2 = IND
    for it in ['int', 'long long']:
        args = array_fact_args(L, it)
        for i in range(r):
            args += [f'({it})boundsLow[{i}]', f'({it})cnt{i}', f'({it})loopSteps[{i}]']
        args += [f'({it})nbx', f'({it})nby', f'({it})nbz', 'nblocks', '(int)tune[DVMH_AMD_T_XCD]']
        args += red_fact_args(L)
        ks = [(c, k) for c, k in kernels if (('rt_INT' in c) == (it == 'int'))]
        b += [ind2 + s for s in launch_dispatch(ks, args, 'dim3((unsigned)nblocks)')]
    if L.stencil:
        b.append(IND + '}')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 1);')
    b.append('}')
    b += handler_reds_finish(L)
    h.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    h.append('}')
    return '\n'.join(h) + '\n'


# ----------------------------------------------------------------- GPU: batched PARALLEL loops
def batch_desc_layout(L):
    """Per-loop descriptor (array of longs): array bases and strides, scalar
    device addresses, (lo, cnt, step) per axis, tile counts, reduction inits."""
    lay, q = {}, 0
    for p in L.arrays:
        lay[('base', p[1])] = q
        q += 1
        lay[('str', p[1])] = q
        q += p[3]
    for p in L.scalars:
        lay[('sc', p[1])] = q
        q += 1
    lay['bounds'] = q
    q += 3 * L.rank
    lay['nb'] = q
    q += 3
    for red in L.reds:
        lay[('ri', red[1])] = q
        q += 1
        if len(red) > 3:
            lay[('ri', red[3])] = q
            q += 1
    return lay, q


def batch_kernel(L):
    """Persistent kernel over all instances' tiles.  The tile body runs in a
    separate inline function whose array bases are __restrict__ parameters
    (distinct DVM arrays never overlap), so once inlined the loads of its
    DEPTH rows may be issued ahead of the stores, as in the per-loop kernels
    whose bases are kernel parameters."""
    r = L.rank
    lay, dlen = batch_desc_layout(L)
    # ---- the tile function
    tp = []
    for p in L.arrays:
        _, nm, ct, ar = p
        tp.append(f'{ct} *__restrict__ {nm}_base')
        tp += [f'IndexT {nm}_s{q}' for q in range(ar)]
    for i in range(r):
        tp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    tp += ['long nbx', 'long nby', 'long nbz', 'long tl']
    for p in L.scalars:
        tp.append(f'{p[2]} &{p[1]}')
    for red in L.reds:
        tp.append(f'{red[2]} &{red[1]}')
        if len(red) > 3:
            tp.append(f'{red[4]} &{red[3]}')
    f = ['template <typename IndexT, int DEPTH>',
         f'__device__ static inline void {L.name}_batch_tile({", ".join(tp)}) {{']
    t = []
    for p in L.arrays:
        _, nm, ct, ar = p
        t.append(f'const IndexT {nm}_str[{ar}] = {{{", ".join(f"{nm}_s{q}" for q in range(ar))}}};')
        t.append(f'const auto {nm} = dvmh_dev::makeView<{ct}, IndexT, {ar}, true>({nm}_base, {nm}_str);')
    for v, ct in L.vars:
        t.append(f'{ct} {v};')
    for pr in L.privates:
        t.append(pr + ';')
    t.append('const IndexT bx = (IndexT)(tl % nbx); tl /= nbx;')
    t.append('const IndexT by = (IndexT)(tl % nby); tl /= nby;')
    t.append('const IndexT bz = (IndexT)(tl % nbz); tl /= nbz;')
    ax_x, ax_y, ax_z = r - 1, (r - 2 if r >= 2 else None), (r - 3 if r >= 3 else None)
    dep_ax = ax_z if ax_z is not None else (ax_y if ax_y is not None else ax_x)
    for a in reversed(range(0, r - 3)):
        t.append(f'const IndexT t{a} = (IndexT)(tl % cnt{a}); tl /= cnt{a};')
    t.append(f'const IndexT t{ax_x}_0 = bx * (IndexT)blockDim.x + (IndexT)threadIdx.x;')
    if ax_y is not None:
        t.append(f'const IndexT t{ax_y}_0 = by * (IndexT)blockDim.y + (IndexT)threadIdx.y;')
    if ax_z is not None:
        t.append(f'const IndexT t{ax_z}_0 = bz * (IndexT)blockDim.z + (IndexT)threadIdx.z;')
    conds = []
    for a in [ax_x, ax_y, ax_z]:
        if a is None:
            continue
        if a == dep_ax and r >= 2:
            t.append(f'const IndexT t{a}_b = t{a}_0 * DEPTH;')
        else:
            t.append(f'const IndexT t{a} = t{a}_0;')
            conds.append(f't{a} < cnt{a}')
    def rows(guarded):
        inner = []
        if r >= 2:
            inner.append(f'const IndexT t{dep_ax} = t{dep_ax}_b + dd;')
        cs = conds + ([f't{dep_ax} < cnt{dep_ax}'] if r >= 2 else [])
        if guarded:
            inner.append(f'if ({" && ".join(cs)}) {{')
        else:
            inner.append('{')
        for i in range(r):
            inner.append(IND + var_assign(L, i, f't{i}', f'lo{i}', f'st{i}'))
        inner.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 1))
        inner.append('}')
        out = ['#pragma unroll', 'for (int dd = 0; dd < DEPTH; dd++) {']
        out += [IND + ln for ln in '\n'.join(inner).split('\n')]
        out.append('}')
        return out
    if r >= 2:
        # a strip wholly inside the bounds runs unguarded: one basic block, so
        # the loads of all DEPTH rows can be issued before the first store
        full = conds + [f't{dep_ax}_b + DEPTH <= cnt{dep_ax}']
        t.append(f'if ({" && ".join(full)}) {{')
        t += [IND + ln for ln in rows(False)]
        t.append('} else {')
        t += [IND + ln for ln in rows(True)]
        t.append('}')
    else:
        t.append(f'if ({" && ".join(conds)}) {{')
        for i in range(r):
            t.append(IND + var_assign(L, i, f't{i}', f'lo{i}', f'st{i}'))
        t.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 1))
        t.append('}')
    f.append(indent('\n'.join(t) + '\n', 1).rstrip('\n'))
    f.append('}')
    # ---- the kernel
    fp = ['const long *__restrict__ desc', 'const long *__restrict__ tileStart', 'long nloops', 'long totalTiles']
    for red in L.reds:
        fp.append(f'{red[2]} *{red[1]}_grid')
        if len(red) > 3:
            fp.append(f'{red[4]} *{red[3]}_grid')
    k = ['template <typename IndexT, int DEPTH>',
         f'__global__ __launch_bounds__(1024) void {L.name}_batch_kernel({", ".join(fp)}) {{']
    b = ['for (long tile = blockIdx.x; tile < totalTiles; tile += gridDim.x) {']
    t = []
    t.append('long lo_ = 0, hi_ = nloops - 1;')
    t.append('while (lo_ < hi_) { long mid = (lo_ + hi_ + 1) >> 1; if (tileStart[mid] <= tile) lo_ = mid; else hi_ = mid - 1; }')
    t.append(f'const long *d = desc + lo_ * {dlen};')
    args = []
    for p in L.arrays:
        _, nm, ct, ar = p
        so = lay[('str', nm)]
        args.append(f'dvmh_dev::globalPtr<{ct}>(d[{lay[("base", nm)]}])')
        args += [f'(IndexT)d[{so + q}]' for q in range(ar)]
    bo = lay['bounds']
    for i in range(r):
        args += [f'(IndexT)d[{bo + 3 * i}]', f'(IndexT)d[{bo + 3 * i + 1}]', f'(IndexT)d[{bo + 3 * i + 2}]']
    nbo = lay['nb']
    args += [f'd[{nbo}]', f'd[{nbo + 1}]', f'd[{nbo + 2}]', 'tile - tileStart[lo_]']
    for p in L.scalars:
        args.append(f'*dvmh_dev::globalPtr<{p[2]}>(d[{lay[("sc", p[1])]}])')
    for red in L.reds:
        t.append(f'{red[2]} {red[1]} = *(const {red[2]} *)&d[{lay[("ri", red[1])]}];')
        args.append(red[1])
        if len(red) > 3:
            t.append(f'{red[4]} {red[3]} = *(const {red[4]} *)&d[{lay[("ri", red[3])]}];')
            args.append(red[3])
    t.append(f'{L.name}_batch_tile<IndexT, DEPTH>({", ".join(args)});')
    t += red_kernel_tail(L, 'tile', accumulate=False)
    b += [IND + ln for ln in '\n'.join(t).split('\n')]
    b.append('}')
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    return '\n'.join(f) + '\n\n' + '\n'.join(k) + '\n'


def batch_march_kernel(L):
    """Persistent kernel over all instances' march tiles of a stencil loop:
    each tile (a strip of the inner axes times `march` planes of the outermost)
    runs the LDS march of the per-loop march kernel (march_tile)."""
    r = L.rank
    lay, dlen = batch_desc_layout(L)
    fp = ['const long *__restrict__ desc', 'const long *__restrict__ tileStart', 'long nloops', 'long totalTiles', 'IndexT march']
    for red in L.reds:
        fp.append(f'{red[2]} *{red[1]}_grid')
        if len(red) > 3:
            fp.append(f'{red[4]} *{red[3]}_grid')
    k = ['template <typename IndexT, int COLS, int NPF>',
         f'__global__ __launch_bounds__(1024) void {L.name}_batch_march_kernel({", ".join(fp)}) {{']
    b = ['for (long tile = blockIdx.x; tile < totalTiles; tile += gridDim.x) {']
    t = ['long lo_ = 0, hi_ = nloops - 1;',
         'while (lo_ < hi_) { long mid = (lo_ + hi_ + 1) >> 1; if (tileStart[mid] <= tile) lo_ = mid; else hi_ = mid - 1; }',
         f'const long *d = desc + lo_ * {dlen};']
    args = []
    for p in L.arrays:
        _, nm, ct, ar = p
        so = lay[('str', nm)]
        args.append(f'dvmh_dev::globalPtr<{ct}>(d[{lay[("base", nm)]}])')
        args += [f'(IndexT)d[{so + q}]' for q in range(ar)]
    for p in L.scalars:
        args.append(f'dvmh_dev::globalPtr<{p[2]}>(d[{lay[("sc", p[1])]}])')
    bo = lay['bounds']
    for i in range(r):
        args += [f'(IndexT)d[{bo + 3 * i}]', f'(IndexT)d[{bo + 3 * i + 1}]', f'(IndexT)d[{bo + 3 * i + 2}]']
    nbo = lay['nb']
    args += [f'(IndexT)d[{nbo}]', f'(IndexT)d[{nbo + 1}]', f'(IndexT)d[{nbo + 2}]', 'march', 'tile - tileStart[lo_]', '(IndexT)0']
    for red in L.reds:
        t.append(f'{red[2]} {red[1]} = *(const {red[2]} *)&d[{lay[("ri", red[1])]}];')
        args.append(red[1])
        if len(red) > 3:
            t.append(f'{red[4]} {red[3]} = *(const {red[4]} *)&d[{lay[("ri", red[3])]}];')
            args.append(red[3])
    t.append(f'{L.name}_march_tile<IndexT, COLS, NPF>({", ".join(args)});')
    t += red_kernel_tail(L, 'tile', accumulate=False)
    b += [IND + ln for ln in '\n'.join(t).split('\n')]
    b.append('}')
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    return '\n'.join(k) + '\n'


def batch_handler(L):
    r = L.rank
    lay, dlen = batch_desc_layout(L)
    h = [f'extern "C" void {L.name}_cuda_batch(DvmType *pBatchRef) {{']
    b = ['DvmType batchRef = *pBatchRef;',
         'long nl = (long)dvmh_amd_batch_count_C(batchRef);',
         'DvmType lr0 = dvmh_amd_batch_loop_C(batchRef, 0);',
         'DvmType tune[DVMH_AMD_TUNE_N];',
         'dvmh_amd_loop_get_tuning_C(lr0, DVMH_AMD_KIND_BATCH, tune);',
         'dim3 threads(0, 0, 0);',
         'hipStream_t stream;',
         *regs_query(f'{L.name}_batch_kernel<int, 1>', 'lr0'),
         'int depth = (int)tune[DVMH_AMD_T_DEPTH];',
         'if (depth != 1 && depth != 2 && depth != 4 && depth != 8) depth = 1;',
         'int bvariant = 0, mcols = 1, npf = 0;',
         'DvmType march = 1;',
         'size_t ldsBytes = 0;']
    bm = L.stencil and r in (2, 3)
    if bm:
        # LDS march tiles (VARIANT 1): same shape rules as the per-loop handler
        shape_ok = 'threads.z == 1' + (' && threads.y == 1' if r == 2 else '')
        rows = '((size_t)threads.y + 2)' if r == 3 else '(size_t)1'
        b += ['bvariant = (int)tune[DVMH_AMD_T_VARIANT];',
              'mcols = (int)tune[DVMH_AMD_T_TILE_A];',
              'march = tune[DVMH_AMD_T_CHUNK];']
        if r == 2:
            # the default batch configuration leaves a stencil batch its own shape:
            # march tiles of 256 threads x 4 points, 8 rows each
            b += ['if (tune[DVMH_AMD_T_OWN] == 1) {',
                  IND + 'bvariant = 1;',
                  IND + 'mcols = 4;',
                  IND + 'march = 8;',
                  IND + 'threads = dim3(256, 1, 1);',
                  '}']
        b += [
              f'if (bvariant == 1 && {shape_ok} && mcols >= 1 && mcols <= 4 && march >= 1) {{',
              IND + f'size_t mplane = {rows} * ((size_t)threads.x * mcols + 2);']
        for nm in L.stencil:
            ct = [p for p in L.arrays if p[1] == nm][0][2]
            b.append(IND + f'ldsBytes += (3 * mplane * sizeof({ct}) + 15) & ~(size_t)15;')
        b += [IND + 'size_t mthreads = (size_t)threads.x * threads.y;',
              IND + 'size_t need = (mplane + mthreads - 1) / mthreads;']
        for cols, sets in MARCH_NPF.items():
            b.append(IND + f'if (mcols == {cols}) {{ ' + ' '.join(f'if (npf == 0 && need <= {v}) npf = {v};' for v in sets) + ' }')
        b += [IND + 'if (npf == 0 || ldsBytes > 64512) bvariant = 0;',
              '} else {',
              IND + 'bvariant = 0;',
              '}']
    b += ['int nonunit = 0;',
          f'long *desc = (long *)calloc((size_t)nl * {dlen}, sizeof(long));',
         'DvmType *tstart = (DvmType *)calloc((size_t)nl + 1, sizeof(DvmType));',
         'long tiles = 0;',
         'int wide = 0;',
         'for (long li = 0; li < nl; li++) {']
    lb = ['DvmType lr = dvmh_amd_batch_loop_C(batchRef, li);',
          f'long *d = desc + li * {dlen};',
          'long maxStride = 1;']
    for pi, p in enumerate(L.params):
        if p[0] == 'array':
            _, nm, ct, ar = p
            lb.append(f'{{ DvmType *hdr = (DvmType *)dvmh_amd_batch_param_C(batchRef, li, {pi});')
            lb.append(f'  {ct} *base = ({ct} *)dvmh_get_natural_base_C(1, hdr);')
            lb.append('  DvmType devHdr[64 + 7];')
            lb.append('  if (dvmh_fill_header_C(1, base, hdr, devHdr, 0) != 0) { fprintf(stderr, "batched loop: array not in natural layout\\n"); abort(); }')
            lb.append(f'  d[{lay[("base", nm)]}] = (long)base;')
            for q in range(ar):
                lb.append(f'  d[{lay[("str", nm)] + q}] = devHdr[{q + 1}];')
            lb.append(f'  for (int q = 0; q < {ar}; q++) if (labs(devHdr[q + 1]) > maxStride) maxStride = labs(devHdr[q + 1]); }}')
        else:
            lb.append(f'd[{lay[("sc", p[1])]}] = (long)dvmh_get_device_addr_C(1, dvmh_amd_batch_param_C(batchRef, li, {pi}));')
    lb.append(f'DvmType blo[{r}], bhi[{r}], bst[{r}];')
    lb.append('dvmh_loop_fill_bounds_C(lr, blo, bhi, bst);')
    for i in range(r):
        lb.append(f'long cnt{i} = (bhi[{i}] - blo[{i}]) / bst[{i}] + 1;')
        lb.append(f'd[{lay["bounds"] + 3 * i}] = blo[{i}]; d[{lay["bounds"] + 3 * i + 1}] = cnt{i}; d[{lay["bounds"] + 3 * i + 2}] = bst[{i}];')
    ax_x, ax_y, ax_z = r - 1, (r - 2 if r >= 2 else None), (r - 3 if r >= 3 else None)
    dep_ax = ax_z if ax_z is not None else (ax_y if ax_y is not None else ax_x)
    lb.append(f'long nbx = (cnt{ax_x} + threads.x - 1) / threads.x;')
    if ax_y is not None:
        dy = ' * depth' if dep_ax == ax_y else ''
        lb.append(f'long nby = (cnt{ax_y} + threads.y{dy} - 1) / (threads.y{dy});')
    else:
        lb.append('long nby = 1;')
    if ax_z is not None:
        lb.append(f'long nbz = (cnt{ax_z} + threads.z * depth - 1) / (threads.z * depth);')
    else:
        lb.append('long nbz = 1;')
    if bm:
        lb.append('if (bvariant == 1) {')
        lb.append(IND + 'if (' + ' || '.join(f'bst[{i}] != 1' for i in range(r)) + ') nonunit = 1;')
        lb.append(IND + f'nbx = (cnt{ax_x} + (long)threads.x * mcols - 1) / ((long)threads.x * mcols);')
        lb.append(IND + (f'nby = (cnt{ax_y} + threads.y - 1) / threads.y;' if r == 3 else 'nby = 1;'))
        lb.append(IND + 'nbz = (cnt0 + march - 1) / march;')
        lb.append('}')
    lb.append(f'd[{lay["nb"]}] = nbx; d[{lay["nb"] + 1}] = nby; d[{lay["nb"] + 2}] = nbz;')
    for i in range(r):
        lb.append(f'if ((labs(blo[{i}]) + cnt{i} * labs(bst[{i}]) + 64) * maxStride > 2147483647L) wide = 1;')
    lb.append('if (dvmh_loop_guess_index_type_C(lr) != rt_INT) wide = 1;')
    for ri, red in enumerate(L.reds):
        lb.append(f'{{ {red[2]} v; ')
        if len(red) > 3:
            lb.append(f'  {red[4]} lv; dvmh_loop_red_init_C(lr, {ri + 1}, &v, &lv);')
            lb.append(f'  memcpy(&d[{lay[("ri", red[3])]}], &lv, sizeof(lv));')
        else:
            lb.append(f'  dvmh_loop_red_init_C(lr, {ri + 1}, &v, 0);')
        lb.append(f'  memcpy(&d[{lay[("ri", red[1])]}], &v, sizeof(v)); }}')
    lb.append('tstart[li] = tiles;')
    lb.append('tiles += nbx * nby * nbz' + ''.join(f' * cnt{a}' for a in range(0, r - 3)) + ';')
    b += [IND + s for s in lb]
    b.append('}')
    b.append('tstart[nl] = tiles;')
    if bm:
        # a loop with non-unit steps, or 64-bit indices: every loop of the batch runs the direct tiles
        b.append('if (bvariant == 1 && (nonunit || wide)) {')
        b.append(IND + 'bvariant = 0;')
        b.append(IND + 'tiles = 0;')
        b.append(IND + 'for (long li = 0; li < nl; li++) {')
        b.append(IND * 2 + f'long *d = desc + li * {dlen};')
        for i in range(r):
            b.append(IND * 2 + f'long cnt{i} = d[{lay["bounds"] + 3 * i + 1}];')
        b.append(IND * 2 + f'long nbx = (cnt{ax_x} + threads.x - 1) / threads.x;')
        if ax_y is not None:
            dy = ' * depth' if dep_ax == ax_y else ''
            b.append(IND * 2 + f'long nby = (cnt{ax_y} + threads.y{dy} - 1) / (threads.y{dy});')
        else:
            b.append(IND * 2 + 'long nby = 1;')
        if ax_z is not None:
            b.append(IND * 2 + f'long nbz = (cnt{ax_z} + threads.z * depth - 1) / (threads.z * depth);')
        else:
            b.append(IND * 2 + 'long nbz = 1;')
        b.append(IND * 2 + f'd[{lay["nb"]}] = nbx; d[{lay["nb"] + 1}] = nby; d[{lay["nb"] + 2}] = nbz;')
        b.append(IND * 2 + 'tstart[li] = tiles;')
        b.append(IND * 2 + 'tiles += nbx * nby * nbz' + ''.join(f' * cnt{a}' for a in range(0, r - 3)) + ';')
        b.append(IND + '}')
        b.append(IND + 'tstart[nl] = tiles;')
        b.append('}')
    b.append('// descriptors and tile starts usually repeat from one time step to the next:')
    b.append('// the runtime keeps their device copies and uploads only what changed')
    b.append('static const char dvmh_tag = 0;')
    b.append(f'long *dDesc = (long *)dvmh_amd_batch_persist_C(batchRef, &dvmh_tag, 0, desc, nl * {dlen} * (long)sizeof(long));')
    b.append('long *dStart = (long *)dvmh_amd_batch_persist_C(batchRef, &dvmh_tag, 1, tstart, (nl + 1) * (long)sizeof(long));')
    for red in L.reds:
        b.append(f'{red[2]} *{red[1]}_grid = ({red[2]} *)dvmh_amd_batch_scratch_C(batchRef, (tiles > 0 ? tiles : 1) * (long)sizeof({red[2]}));')
        if len(red) > 3:
            b.append(f'{red[4]} *{red[3]}_grid = ({red[4]} *)dvmh_amd_batch_scratch_C(batchRef, (tiles > 0 ? tiles : 1) * (long)sizeof({red[4]}));')
    b.append('long wgs = tiles < 16384 ? tiles : 16384;   // persistent: each workgroup walks tiles')
    b.append('if (tiles > 0) {')
    b.append(IND + 'dvmh_amd_mark_kernel_C(lr0, 0);')
    ra = []
    for red in L.reds:
        ra.append(f'{red[1]}_grid')
        if len(red) > 3:
            ra.append(f'{red[3]}_grid')
    if bm:
        b.append(IND + 'if (bvariant == 1) {')
        margs = ['dDesc', 'dStart', 'nl', 'tiles', '(int)march'] + ra
        for cols, sets in MARCH_NPF.items():
            for v in sets:
                b.append(IND * 2 + f'if (mcols == {cols} && npf == {v}) hipLaunchKernelGGL(({L.name}_batch_march_kernel<int, {cols}, {v}>), '
                         f'dim3((unsigned)wgs), threads, ldsBytes, stream, {", ".join(margs)});')
        b.append(IND + '} else {')
    pre = IND * 2 if bm else IND
    for it in ['int', 'long long']:
        for dp in [1, 2, 4, 8]:
            cond = f'{"!wide" if it == "int" else "wide"} && depth == {dp}'
            args = ['dDesc', 'dStart', 'nl', 'tiles'] + ra
            b.append(pre + f'if ({cond}) hipLaunchKernelGGL(({L.name}_batch_kernel<{it}, {dp}>), dim3((unsigned)wgs), threads, 0, stream, {", ".join(args)});')
    if bm:
        b.append(IND + '}')
    b.append(IND + 'dvmh_amd_mark_kernel_C(lr0, 1);')
    b.append('}')
    for ri, red in enumerate(L.reds):
        loc = f'{red[3]}_grid' if len(red) > 3 else '0'
        b.append(f'dvmh_amd_batch_red_finish_C(batchRef, {ri + 1}, {red[1]}_grid, {loc}, tstart);')
    b.append('free(desc);')
    b.append('free(tstart);')
    h.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    h.append('}')
    return '\n'.join(h) + '\n'


# ----------------------------------------------------------------- GPU: fused loop pairs
# A PARALLEL stencil loop L1 directly followed (same region, same bounds) by a
# PARALLEL loop L2 that touches every array only at the loop point -- a time
# step's update loop and its copy-back / residual loop -- can run as one
# pass: each point runs L1's body and then L2's.  L1 reads its stencil arrays
# through the LDS march, loaded before anything is written; L2 may overwrite
# such an array at the point only where no other tile reads it, i.e. off the
# tile edges that face another tile.  Points on those edges run L2 in a second
# kernel after the first has finished (every read of the old values is done
# then).  Sequential semantics are kept: L2 at point p sees L1's results at p
# (same thread), and every read L1 makes of an array L2 writes sees the value
# from before L2.

WRITE_OP = r'(?:=(?!=)|\+=|-=|\*=|/=|%=|&=|\|=|\^=|<<=|>>=|\+\+|--)'
# edge points per thread of the border kernel (a workgroup of 256 threads
# takes 256 * PAIR_EDGE_PTS consecutive edge points, thread t every 256th)
PAIR_EDGE_PTS = 8


def array_refs(L, name):
    """Offsets (per loop axis) and write flag of every access name[...] in L's
    body, or None when some access is not name[v0 + c0][v1 + c1]... with the
    loop's own variables in order, or the name is used any other way."""
    body = L.body
    if re.search(r'\b' + name + r'\b(?!\s*\[)', body) or re.search(r'(\+\+|--)\s*' + name + r'\b', body):
        return None
    refs = []
    for m in re.finditer(r'\b' + name + r'\s*((?:\[[^\[\]]*\]\s*)+)(\s*' + WRITE_OP + r')?', body):
        idx = re.findall(r'\[([^\[\]]*)\]', m.group(1))
        if len(idx) != L.rank:
            return None
        offs = []
        for (v, _), e in zip(L.vars, idx):
            mm = re.fullmatch(r'\s*' + re.escape(v) + r'\s*(?:([+-])\s*(\d+)\s*)?', e)
            if not mm:
                return None
            offs.append(0 if not mm.group(1) else int(mm.group(2)) * (1 if mm.group(1) == '+' else -1))
        refs.append((offs, bool(m.group(2))))
    # every name[ must have been matched (nested brackets are not)
    if len(re.findall(r'\b' + name + r'\s*\[', body)) != len(refs):
        return None
    return refs


def pair_eligible(L1, L2):
    """Whether L2 may run fused after L1 (see above): both PARALLEL on the GPU
    with the same rank (2 or 3), L1 with stencil arrays it only reads, every
    other array access of L1 and every array access of L2 at offset 0, scalar
    reductions whose variables L1 does not name, and names that mean the same
    kind of object in both loops.  The pair handler checks at run time that
    equally named arrays are the same array and differently named ones are
    not, and that the bounds agree."""
    for L in (L1, L2):
        if L.across or L.sequential or L.host_only or L.rank not in (2, 3):
            return False
        if any(p[0] not in ('array', 'scalar') for p in L.params):
            return False
        if any(L.rlen(red) > 1 for red in L.reds):
            return False
    if not L1.stencil or L1.rank != L2.rank:
        return False
    kinds1 = {p[1]: p for p in L1.params}
    for p in L2.params:
        q = kinds1.get(p[1])
        if q is not None and tuple(q) != tuple(p):
            return False
    for L, stencil in ((L1, L1.stencil), (L2, [])):
        for p in L.arrays:
            if p[3] != L.rank:
                return False
            refs = array_refs(L, p[1])
            if refs is None:
                return False
            if p[1] in stencil:
                continue
            if any(any(o != 0 for o in offs) for offs, _ in refs):
                return False
    names1 = set(kinds1) | {v for v, _ in L1.vars}
    for red in L1.reds:
        names1.add(red[1])
        if len(red) > 3:
            names1.add(red[3])
    for red in L2.reds:
        if red[1] in names1 or (len(red) > 3 and red[3] in names1):
            return False
    return True


def pair_name(L1, L2):
    return f'{L1.name}__{L2.name}'


def pair_tail_params(L1, L2):
    """Extra parameters of the pair's tile function: L2's arrays and scalars
    that L1 does not have, and L2's reduction variables."""
    n1 = {p[1] for p in L1.params}
    tp = []
    for p in L2.arrays:
        _, nm, ct, ar = p
        if nm in n1:
            continue
        tp.append(f'{ct} *__restrict__ p2_{nm}_base')
        tp += [f'IndexT p2_{nm}_s{k}' for k in range(ar)]
    for p in L2.scalars:
        if p[1] not in n1:
            tp.append(f'{p[2]} *__restrict__ p2_{p[1]}_ptr')
    for red in L2.reds:
        tp.append(f'{red[2]} &p2_{red[1]}')
        if len(red) > 3:
            tp.append(f'{red[4]} &p2_{red[3]}')
    return tp


def lds_reads(L1, L2):
    """For each of L1's stencil arrays that L2 reads (at its point), the text
    spans of those reads in L2's body when every one of them comes before the
    first write of the array in the body text and the body has no loop: such a
    read yields the value L1's LDS window holds at the point (the value before
    L2, nothing having written it yet), so it need not go back to memory."""
    if re.search(r'\b(for|while|do|goto)\b', L2.body):
        return {}
    out = {}
    for p in L2.arrays:
        nm = p[1]
        if nm not in L1.stencil:
            continue
        reads, writes = [], []
        for m in re.finditer(r'\b' + nm + r'\s*((?:\[[^\[\]]*\]\s*)+)(\s*' + WRITE_OP + r')?', L2.body):
            if m.group(2):
                writes.append(m.start())
            else:
                reads.append((m.start(), m.end(1)))
        if not reads or (writes and max(r[0] for r in reads) > min(writes)):
            continue
        out[nm] = reads
    return out


def pair_tail_scope(L1, L2, tname):
    """L2's names bound for one point (in a scope of their own, so they may
    shadow L1's), its loop variables from the point's offsets tname(axis), then
    its body.  Arrays L1 also has use L1's base (the same array); reads that
    lds_reads allows take the point's value from L1's LDS window, captured
    before the array's name is rebound."""
    n1 = {p[1] for p in L1.params}
    s = []
    body = L2.body
    spans = []
    for nm, reads in lds_reads(L1, L2).items():
        ct = [p for p in L2.arrays if p[1] == nm][0][2]
        idx = ''.join(f'[(IndexT)(lo{i} + {tname(i)})]' for i in range(L2.rank))
        s.append(f'const {ct} dvmhc_{nm} = {nm}{idx};   // the point\'s value in the LDS window')
        spans += [(a, b, f'dvmhc_{nm} ') for a, b in reads]
    for a, b, t in sorted(spans, reverse=True):
        body = body[:a] + t + body[b:]
    for p in L2.arrays:
        _, nm, ct, ar = p
        pre = '' if nm in n1 else 'p2_'
        s.append(f'const IndexT p2_{nm}_str[{ar}] = {{{", ".join(f"{pre}{nm}_s{k}" for k in range(ar))}}};')
        s.append(f'const auto {nm} = dvmh_dev::makeView<{ct}, IndexT, {ar}, true>({pre}{nm}_base, p2_{nm}_str);')
    for p in L2.scalars:
        pre = '' if p[1] in n1 else 'p2_'
        s.append(scalar_bind(L2.body, p, pre))
    for red in L2.reds:
        s.append(f'{red[2]} &{red[1]} = p2_{red[1]};')
        if len(red) > 3:
            s.append(f'{red[4]} &{red[3]} = p2_{red[3]};')
    for v, ct in L2.vars:
        s.append(f'{ct} {v};')
    for pr in L2.privates:
        s.append(pr + ';')
    for i in range(L2.rank):
        s.append(var_assign(L2, i, tname(i), f'lo{i}', '1'))
    s.append('{\n' + indent(body + '\n', 1) + '}')
    return s


def pair_tail(L1, L2, ax_x, ax_y):
    """Inside the march's point loop: L2 at the point unless it lies on a tile
    edge facing another tile (those run in the pair's border kernel)."""
    edge = ['(d == 0 && bz > 0)', '(d == march - 1 && bz + 1 < nbz)',
            f'(t{ax_x} == tx0 && bx > 0)', f'(t{ax_x} == tx0 + tileX - 1 && bx + 1 < nbx)']
    if ax_y is not None:
        edge += ['(threadIdx.y == 0 && by > 0)', '(threadIdx.y == blockDim.y - 1 && by + 1 < nby)']
    s = ['// the fused loop at this point, unless another tile reads it',
         f'if (!({" || ".join(edge)})) {{']
    s += [IND + ln for ln in '\n'.join(pair_tail_scope(L1, L2, lambda i: f't{i}')).split('\n')]
    s.append('}')
    return s


def pair_desc_layout(L1, L2):
    """Per pair: L1's batch descriptor, then L2's, then the start of the pair's
    segment of L2's reduction slots and its number of march tiles."""
    lay1, n1 = batch_desc_layout(L1)
    lay2, n2 = batch_desc_layout(L2)
    lay2 = {k: v + n1 for k, v in lay2.items()}
    ext = {'seg2': n1 + n2, 'ntiles': n1 + n2 + 1}
    return lay1, lay2, ext, n1 + n2 + 2


class _Renamed:
    """A loop's reductions under prefixed names (for red_kernel_tail)."""
    def __init__(self, L, pre):
        self.reds = [tuple([r[0], pre + r[1], r[2]] + ([pre + r[3], r[4]] if len(r) > 3 else [])) for r in L.reds]

    def rlen(self, red):
        return 1


def pair_kernels(L1, L2):
    """The pair's march kernel (all pairs' tiles, L1 then L2 per interior
    point) and its border kernel (L2 at the tile-edge points)."""
    r = L1.rank
    P = pair_name(L1, L2)
    lay1, lay2, ext, dlen = pair_desc_layout(L1, L2)
    n1 = {p[1] for p in L1.params}
    R2 = _Renamed(L2, 'p2_')
    grids = []
    for red in L1.reds:
        grids.append(f'{red[2]} *{red[1]}_grid')
        if len(red) > 3:
            grids.append(f'{red[4]} *{red[3]}_grid')
    grids2 = []
    for red in R2.reds:
        grids2.append(f'{red[2]} *{red[1]}_grid')
        if len(red) > 3:
            grids2.append(f'{red[4]} *{red[3]}_grid')
    fp = ['const long *__restrict__ desc', 'const long *__restrict__ tileStart', 'long npairs', 'long totalTiles', 'IndexT march',
          'int xcd'] + grids + grids2
    k = ['template <typename IndexT, int COLS, int NPF>',
         f'__global__ __launch_bounds__(1024) void {P}_pair_kernel({", ".join(fp)}) {{']
    b = ['for (long t_ = blockIdx.x; t_ < totalTiles; t_ += gridDim.x) {']
    t = ['const long tile = dvmh_dev::xcdTile(t_, totalTiles, xcd);',
         'long lo_ = 0, hi_ = npairs - 1;',
         'while (lo_ < hi_) { long mid = (lo_ + hi_ + 1) >> 1; if (tileStart[mid] <= tile) lo_ = mid; else hi_ = mid - 1; }',
         f'const long *d = desc + lo_ * {dlen};']
    args = []
    for p in L1.arrays:
        _, nm, ct, ar = p
        so = lay1[('str', nm)]
        args.append(f'dvmh_dev::globalPtr<{ct}>(d[{lay1[("base", nm)]}])')
        args += [f'(IndexT)d[{so + q}]' for q in range(ar)]
    for p in L1.scalars:
        args.append(f'dvmh_dev::globalPtr<{p[2]}>(d[{lay1[("sc", p[1])]}])')
    bo = lay1['bounds']
    for i in range(r):
        args += [f'(IndexT)d[{bo + 3 * i}]', f'(IndexT)d[{bo + 3 * i + 1}]', f'(IndexT)d[{bo + 3 * i + 2}]']
    nbo = lay1['nb']
    args += [f'(IndexT)d[{nbo}]', f'(IndexT)d[{nbo + 1}]', f'(IndexT)d[{nbo + 2}]', 'march', 'tile - tileStart[lo_]', '(IndexT)0']
    for red in L1.reds:
        t.append(f'{red[2]} {red[1]} = *(const {red[2]} *)&d[{lay1[("ri", red[1])]}];')
        args.append(red[1])
        if len(red) > 3:
            t.append(f'{red[4]} {red[3]} = *(const {red[4]} *)&d[{lay1[("ri", red[3])]}];')
            args.append(red[3])
    for p in L2.arrays:
        _, nm, ct, ar = p
        if nm in n1:
            continue
        so = lay2[('str', nm)]
        args.append(f'dvmh_dev::globalPtr<{ct}>(d[{lay2[("base", nm)]}])')
        args += [f'(IndexT)d[{so + q}]' for q in range(ar)]
    for p in L2.scalars:
        if p[1] not in n1:
            args.append(f'dvmh_dev::globalPtr<{p[2]}>(d[{lay2[("sc", p[1])]}])')
    for red in L2.reds:
        t.append(f'{red[2]} p2_{red[1]} = *(const {red[2]} *)&d[{lay2[("ri", red[1])]}];')
        args.append(f'p2_{red[1]}')
        if len(red) > 3:
            t.append(f'{red[4]} p2_{red[3]} = *(const {red[4]} *)&d[{lay2[("ri", red[3])]}];')
            args.append(f'p2_{red[3]}')
    t.append(f'{P}_pair_tile<IndexT, COLS, NPF>({", ".join(args)});')
    t += red_kernel_tail(L1, 'tile', accumulate=False)
    t += red_kernel_tail(R2, f'd[{ext["seg2"]}] + (tile - tileStart[lo_])', accumulate=False)
    b += [IND + ln for ln in '\n'.join(t).split('\n')]
    b.append('}')
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    # ---- border kernel: L2 at the points on tile edges that face another tile.
    # Per axis a (march axis, then y for rank 3, then x) with tile size T and n
    # tiles, the edge indices are the 2 (n - 1) indices next to a boundary
    # between tiles; slab a lists the points whose a-index is one of them and
    # whose earlier axes' indices are not (each edge point once).
    ax = [0, 1, 2] if r == 3 else [0, 1]
    fpb = ['const long *__restrict__ desc', 'const long *__restrict__ blkStart', 'long npairs', 'long totalBlocks', 'IndexT march',
           'IndexT tileY', 'IndexT tileX']
    for red in L2.reds:
        fpb.append(f'{red[2]} *{red[1]}_grid')
        if len(red) > 3:
            fpb.append(f'{red[4]} *{red[3]}_grid')
    g = ['template <typename IndexT>',
         f'__global__ __launch_bounds__(256) void {P}_pair_border_kernel({", ".join(fpb)}) {{']
    gb = ['for (long blk = blockIdx.x; blk < totalBlocks; blk += gridDim.x) {']
    u = ['long lo_ = 0, hi_ = npairs - 1;',
         'while (lo_ < hi_) { long mid = (lo_ + hi_ + 1) >> 1; if (blkStart[mid] <= blk) lo_ = mid; else hi_ = mid - 1; }',
         f'const long *d = desc + lo_ * {dlen};']
    u += red_kernel_locals_from(L2, lay2)
    bo2 = lay2['bounds']
    for i in range(r):
        u.append(f'const IndexT lo{i} = (IndexT)d[{bo2 + 3 * i}], cnt{i} = (IndexT)d[{bo2 + 3 * i + 1}];')
    # tile sizes and counts per axis (the counts are L1's nbx, nby, nbz)
    tsz = {0: 'march', ax[-1]: 'tileX'}
    tcnt = {0: f'd[{nbo + 2}]', ax[-1]: f'd[{nbo}]'}
    if r == 3:
        tsz[1] = 'tileY'
        tcnt[1] = f'd[{nbo + 1}]'
    # (32-bit arithmetic: the handler takes this path only for 32-bit indices)
    for a in ax:
        u.append(f'const IndexT T{a} = (IndexT){tsz[a]}, n{a} = (IndexT){tcnt[a]}, r{a} = n{a} > 1 ? 2 * (n{a} - 1) : 0;')
    u.append('// the index next to boundary u / 2 between tiles: its last (u even) or first (u odd) index')
    u.append('auto edgeAt = [](IndexT uu, IndexT T) -> IndexT { return (uu & 1) ? ((uu >> 1) + 1) * T : ((uu >> 1) + 1) * T - 1; };')
    u.append('auto onEdge = [](IndexT tt, IndexT T, IndexT n) -> bool { const IndexT q = tt / T, m = tt - q * T; return (m == 0 && q > 0) || (m == T - 1 && q + 1 < n); };')
    # L2's arrays and scalars from the descriptor, under names of the kernel's
    # own (bound to L2's names in the point's scope)
    for p in L2.arrays:
        _, nm, ct, ar = p
        so = lay2[('str', nm)]
        u.append(f'{ct} *const dvmhp_{nm} = dvmh_dev::globalPtr<{ct}>(d[{lay2[("base", nm)]}]);')
        u.append(f'const IndexT dvmhp_{nm}_str[{ar}] = {{{", ".join(f"(IndexT)d[{so + q}]" for q in range(ar))}}};')
    for p in L2.scalars:
        u.append(f'{p[2]} *const dvmhp_{p[1]} = dvmh_dev::globalPtr<{p[2]}>(d[{lay2[("sc", p[1])]}]);')
    u.append(f'for (int dvmh_k = 0; dvmh_k < {PAIR_EDGE_PTS}; dvmh_k++) {{')
    pu = [f'const IndexT e = (IndexT)(blk - blkStart[lo_]) * (IndexT)blockDim.x * {PAIR_EDGE_PTS} + (IndexT)dvmh_k * (IndexT)blockDim.x + (IndexT)threadIdx.x;',
          'bool valid = false;']
    pu.append(' '.join(f'IndexT p{a} = 0;' for a in ax))
    ext_cnt = {a: f'cnt{a}' for a in ax}
    start = '0'
    for si, a in enumerate(ax):
        dims = [(f'r{a}' if b == a else ext_cnt[b]) for b in ax]
        size = ' * '.join(dims)
        cond = 'if' if si == 0 else 'else if'
        pu.append(f'{cond} (e < {start} + {size}) {{')
        pu.append(IND + f'IndexT f = e - ({start});')
        # decode with the innermost axis fastest
        for bi in reversed(range(len(ax))):
            b_ = ax[bi]
            dim = dims[bi]
            if bi > 0:
                pu.append(IND + f'const IndexT c{b_} = f % ({dim}); f /= ({dim});')
            else:
                pu.append(IND + f'const IndexT c{b_} = f;')
        for b_ in ax:
            pu.append(IND + (f'p{b_} = edgeAt(c{b_}, T{b_});' if b_ == a else f'p{b_} = c{b_};'))
        earlier = [f'!onEdge(p{b_}, T{b_}, n{b_})' for b_ in ax[:si]]
        pu.append(IND + f'valid = {" && ".join(earlier) if earlier else "true"};')
        pu.append('}')
        start = f'{start} + {size}' if start != '0' else size
    pu.append('if (valid) {')
    for a in ax:
        pu.append(IND + f'const IndexT t{a} = p{a};')
    sc = []
    for p in L2.arrays:
        _, nm, ct, ar = p
        sc.append(f'const auto {nm} = dvmh_dev::makeView<{ct}, IndexT, {ar}, true>(dvmhp_{nm}, dvmhp_{nm}_str);')
    for p in L2.scalars:
        sc.append(scalar_bind(L2.body, p, ptr=f'dvmhp_{p[1]}'))
    for v, ct in L2.vars:
        sc.append(f'{ct} {v};')
    for pr in L2.privates:
        sc.append(pr + ';')
    for i in range(r):
        sc.append(var_assign(L2, i, f't{i}', f'lo{i}', '1'))
    sc.append('{\n' + indent(L2.body + '\n', 1) + '}')
    pu += [IND + ln for ln in '\n'.join(sc).split('\n')]
    pu.append('}')
    u += [IND + ln for ln in '\n'.join(pu).split('\n')]
    u.append('}')
    u += red_kernel_tail(L2, f'd[{ext["seg2"]}] + d[{ext["ntiles"]}] + (blk - blkStart[lo_])', accumulate=False)
    gb += [IND + ln for ln in '\n'.join(u).split('\n')]
    gb.append('}')
    g.append(indent('\n'.join(gb) + '\n', 1).rstrip('\n'))
    g.append('}')
    return '\n'.join(k) + '\n\n' + '\n'.join(g) + '\n'


def red_kernel_locals_from(L, lay):
    """Reduction variables of L initialised from a descriptor's init words."""
    s = []
    for red in L.reds:
        s.append(f'{red[2]} {red[1]} = *(const {red[2]} *)&d[{lay[("ri", red[1])]}];')
        if len(red) > 3:
            s.append(f'{red[4]} {red[3]} = *(const {red[4]} *)&d[{lay[("ri", red[3])]}];')
    return s


def pair_border_count(r):
    """Host text: `edges`, the edge points of one pair (numbered as the border
    kernel numbers them), from the extents cnt* and the tile counts nb*."""
    ax = [0, 1, 2] if r == 3 else [0, 1]
    tcnt = {0: 'nbz', ax[-1]: 'nbx'}
    if r == 3:
        tcnt[1] = 'nby'
    s = []
    for a in ax:
        s.append(f'long re{a} = {tcnt[a]} > 1 ? 2 * ({tcnt[a]} - 1) : 0;')
    terms = []
    for a in ax:
        terms.append(' * '.join(f're{b}' if b == a else f'cnt{b}' for b in ax))
    s.append(f'long edges = {" + ".join(terms)};')
    return s


def pair_own_shape(r, lay, dlen):
    """Handler lines: with DVMH_AMD_T_OWN (the default configuration) a rank-2
    pair takes the widest march shape that still gives every CU four tiles
    over all pairs of the call (one 2048^2 grid and 512 grids of 64..1024 need
    different shapes)."""
    if r != 2:
        return []
    c0, c1 = lay['bounds'] + 1, lay['bounds'] + 4
    return [IND + 'if (tune[DVMH_AMD_T_OWN] == 1) {',
            IND * 2 + '// (threads, points per thread, planes), widest first',
            IND * 2 + 'static const int ladder[5][3] = {{256, 4, 32}, {256, 2, 32}, {128, 2, 16}, {64, 2, 16}, {64, 1, 8}};',
            IND * 2 + 'int q = 0;',
            IND * 2 + 'for (; q < 4; q++) {',
            IND * 3 + 'long t = 0;',
            IND * 3 + f'for (long li = 0; li < np; li++) {{ const long *d = desc + li * {dlen}; '
                      f'const long w = (long)ladder[q][0] * ladder[q][1]; t += ((d[{c1}] + w - 1) / w) * ((d[{c0}] + ladder[q][2] - 1) / ladder[q][2]); }}',
            IND * 3 + 'if (t >= 1024) break;',
            IND * 2 + '}',
            IND * 2 + 'threads = dim3((unsigned)ladder[q][0], 1, 1);',
            IND * 2 + 'mcols = ladder[q][1];',
            IND * 2 + 'march = ladder[q][2];',
            IND + '}']


def pair_handler(L1, L2):
    r = L1.rank
    P = pair_name(L1, L2)
    lay1, lay2, ext, dlen = pair_desc_layout(L1, L2)
    a1 = {p[1]: pi for pi, p in enumerate(L1.params) if p[0] == 'array'}
    a2 = {p[1]: pi for pi, p in enumerate(L2.params) if p[0] == 'array'}
    h = [f'extern "C" void {P}_cuda_pair(DvmType *pBatchRef) {{']
    b = ['DvmType batchRef = *pBatchRef;',
         'long np = (long)dvmh_amd_batch_count_C(batchRef);',
         'DvmType lr0 = dvmh_amd_pair_loop_C(batchRef, 0, 0);',
         f'long *desc = (long *)calloc((size_t)np * {dlen}, sizeof(long));',
         'DvmType *tstart = (DvmType *)calloc((size_t)np + 1, sizeof(DvmType));',
         'DvmType *bstart = (DvmType *)calloc((size_t)np + 1, sizeof(DvmType));',
         'DvmType *seg2 = (DvmType *)calloc((size_t)np + 1, sizeof(DvmType));',
         'bool fused = true;',
         '// every pair: the same arrays under the same names, natural layouts,',
         '// equal unit-step bounds and 32-bit indices',
         'for (long li = 0; li < np && fused; li++) {']
    lb = ['DvmType lr1 = dvmh_amd_pair_loop_C(batchRef, 0, li), lr2 = dvmh_amd_pair_loop_C(batchRef, 1, li);',
          f'long *d = desc + li * {dlen};',
          'long maxStride = 1;']
    for n1, p1 in a1.items():
        for n2, p2 in a2.items():
            op = '!=' if n1 == n2 else '=='
            lb.append(f'if (dvmh_amd_pair_param_C(batchRef, 0, li, {p1}) {op} dvmh_amd_pair_param_C(batchRef, 1, li, {p2})) fused = false;')
    for mem, L, lay in ((0, L1, lay1), (1, L2, lay2)):
        for pi, p in enumerate(L.params):
            if p[0] == 'array':
                _, nm, ct, ar = p
                lb.append(f'{{ DvmType *hdr = (DvmType *)dvmh_amd_pair_param_C(batchRef, {mem}, li, {pi});')
                lb.append(f'  void *base = dvmh_get_natural_base_C(1, hdr);')
                lb.append('  DvmType devHdr[64 + 7];')
                lb.append('  if (dvmh_fill_header_C(1, base, hdr, devHdr, 0) != 0) fused = false;')
                lb.append(f'  d[{lay[("base", nm)]}] = (long)base;')
                for q in range(ar):
                    lb.append(f'  d[{lay[("str", nm)] + q}] = devHdr[{q + 1}];')
                lb.append(f'  for (int q = 0; q < {ar}; q++) if (labs(devHdr[q + 1]) > maxStride) maxStride = labs(devHdr[q + 1]); }}')
            else:
                lb.append(f'd[{lay[("sc", p[1])]}] = (long)dvmh_get_device_addr_C(1, dvmh_amd_pair_param_C(batchRef, {mem}, li, {pi}));')
    lb += [f'DvmType blo[{r}], bhi[{r}], bst[{r}], blo2[{r}], bhi2[{r}], bst2[{r}];',
           'dvmh_loop_fill_bounds_C(lr1, blo, bhi, bst);',
           'dvmh_loop_fill_bounds_C(lr2, blo2, bhi2, bst2);',
           f'for (int a = 0; a < {r}; a++) if (bst[a] != 1 || bst2[a] != 1 || blo[a] != blo2[a] || bhi[a] != bhi2[a] || bhi[a] < blo[a]) fused = false;',
           'if (dvmh_loop_guess_index_type_C(lr1) != rt_INT || dvmh_loop_guess_index_type_C(lr2) != rt_INT) fused = false;']
    for i in range(r):
        lb.append(f'long cnt{i} = bhi[{i}] - blo[{i}] + 1;')
        for lay in (lay1, lay2):
            lb.append(f'd[{lay["bounds"] + 3 * i}] = blo[{i}]; d[{lay["bounds"] + 3 * i + 1}] = cnt{i}; d[{lay["bounds"] + 3 * i + 2}] = 1;')
        lb.append(f'if ((labs(blo[{i}]) + cnt{i} + 64) * maxStride > 2147483647L) fused = false;')
    for mem, L, lay in ((0, L1, lay1), (1, L2, lay2)):
        for ri, red in enumerate(L.reds):
            lb.append(f'{{ {red[2]} v; ')
            if len(red) > 3:
                lb.append(f'  {red[4]} lv; dvmh_loop_red_init_C(lr{mem + 1}, {ri + 1}, &v, &lv);')
                lb.append(f'  memcpy(&d[{lay[("ri", red[3])]}], &lv, sizeof(lv));')
            else:
                lb.append(f'  dvmh_loop_red_init_C(lr{mem + 1}, {ri + 1}, &v, 0);')
            lb.append(f'  memcpy(&d[{lay[("ri", red[1])]}], &v, sizeof(v)); }}')
    b += [IND + s for s in lb]
    b.append('}')
    # the pair's launch shape (its own tuner site)
    shape_ok = 'threads.z == 1 && ' + ('threads.y == 1' if r == 2 else 'threads.y >= 2')
    rows = '((size_t)threads.y + 2)' if r == 3 else '(size_t)1'
    b += ['DvmType tune[DVMH_AMD_TUNE_N];',
          'dim3 threads(0, 0, 0);',
          'hipStream_t stream;',
          'int mcols = 0, npf = 0;',
          'long march = 0;',
          'size_t ldsBytes = 0;',
          'if (fused) {',
          IND + 'dvmh_amd_loop_get_tuning_C(lr0, DVMH_AMD_KIND_PAIR, tune);',
          *[IND + x for x in regs_query(f'{P}_pair_kernel<int, 2, 3>', 'lr0')],
          IND + 'mcols = (int)tune[DVMH_AMD_T_TILE_A];',
          IND + 'march = (long)tune[DVMH_AMD_T_CHUNK];',
          *pair_own_shape(r, lay1, dlen),
          IND + f'if (tune[DVMH_AMD_T_VARIANT] != 1 || !({shape_ok}) || mcols < 1 || mcols > 4 || march < 2) fused = false;',
          '}',
          'if (fused) {',
          IND + f'size_t mplane = {rows} * ((size_t)threads.x * mcols + 2);']
    for nm in L1.stencil:
        ct = [p for p in L1.arrays if p[1] == nm][0][2]
        b.append(IND + f'ldsBytes += (3 * mplane * sizeof({ct}) + 15) & ~(size_t)15;')
    b += [IND + 'size_t mthreads = (size_t)threads.x * threads.y;',
          IND + 'size_t need = (mplane + mthreads - 1) / mthreads;']
    for cols, sets in MARCH_NPF.items():
        b.append(IND + f'if (mcols == {cols}) {{ ' + ' '.join(f'if (npf == 0 && need <= {v}) npf = {v};' for v in sets) + ' }')
    b += [IND + 'if (npf == 0 || ldsBytes > 64512) fused = false;',
          '}']
    # otherwise the runtime runs the loops one by one through their own handlers
    b.append('if (!fused) {')
    b.append(IND + 'free(desc); free(tstart); free(bstart); free(seg2);')
    b.append(IND + 'dvmh_amd_pair_decline_C(batchRef);')
    b.append(IND + 'return;')
    b.append('}')
    # tiles and edge blocks per pair
    ax_x = r - 1
    b += ['long tiles = 0, blocks = 0;',
          'for (long li = 0; li < np; li++) {',
          IND + f'long *d = desc + li * {dlen};']
    for i in range(r):
        b.append(IND + f'long cnt{i} = d[{lay1["bounds"] + 3 * i + 1}];')
    b.append(IND + f'long nbx = (cnt{ax_x} + (long)threads.x * mcols - 1) / ((long)threads.x * mcols);')
    b.append(IND + ('long nby = (cnt1 + threads.y - 1) / threads.y;' if r == 3 else 'long nby = 1;'))
    b.append(IND + 'long nbz = (cnt0 + march - 1) / march;')
    for lay in (lay1, lay2):
        b.append(IND + f'd[{lay["nb"]}] = nbx; d[{lay["nb"] + 1}] = nby; d[{lay["nb"] + 2}] = nbz;')
    b += [IND + s for s in pair_border_count(r)]
    b += [IND + '// edge points are numbered in 32-bit arithmetic in the border kernel',
          IND + f'if (edges + {256 * PAIR_EDGE_PTS} > 2147483647L) fused = false;',
          IND + f'long ntl = nbx * nby * nbz, nblk = (edges + {256 * PAIR_EDGE_PTS - 1}) / {256 * PAIR_EDGE_PTS};',
          IND + 'tstart[li] = tiles; bstart[li] = blocks; seg2[li] = tiles + blocks;',
          IND + f'd[{ext["seg2"]}] = tiles + blocks; d[{ext["ntiles"]}] = ntl;',
          IND + 'tiles += ntl; blocks += nblk;',
          '}',
          'tstart[np] = tiles; bstart[np] = blocks; seg2[np] = tiles + blocks;',
          'if (!fused) {',
          IND + 'free(desc); free(tstart); free(bstart); free(seg2);',
          IND + 'dvmh_amd_pair_decline_C(batchRef);',
          IND + 'return;',
          '}',
          'static const char dvmh_tag = 0;',
          f'long *dDesc = (long *)dvmh_amd_batch_persist_C(batchRef, &dvmh_tag, 0, desc, np * {dlen} * (long)sizeof(long));',
          'long *dStart = (long *)dvmh_amd_batch_persist_C(batchRef, &dvmh_tag, 1, tstart, (np + 1) * (long)sizeof(long));',
          'long *dBlk = (long *)dvmh_amd_batch_persist_C(batchRef, &dvmh_tag, 2, bstart, (np + 1) * (long)sizeof(long));']
    ra1, ra2 = [], []
    for red in L1.reds:
        b.append(f'{red[2]} *{red[1]}_grid = ({red[2]} *)dvmh_amd_batch_scratch_C(batchRef, (tiles > 0 ? tiles : 1) * (long)sizeof({red[2]}));')
        ra1.append(f'{red[1]}_grid')
        if len(red) > 3:
            b.append(f'{red[4]} *{red[3]}_grid = ({red[4]} *)dvmh_amd_batch_scratch_C(batchRef, (tiles > 0 ? tiles : 1) * (long)sizeof({red[4]}));')
            ra1.append(f'{red[3]}_grid')
    for red in L2.reds:
        b.append(f'{red[2]} *p2_{red[1]}_grid = ({red[2]} *)dvmh_amd_batch_scratch_C(batchRef, (tiles + blocks > 0 ? tiles + blocks : 1) * (long)sizeof({red[2]}));')
        ra2.append(f'p2_{red[1]}_grid')
        if len(red) > 3:
            b.append(f'{red[4]} *p2_{red[3]}_grid = ({red[4]} *)dvmh_amd_batch_scratch_C(batchRef, (tiles + blocks > 0 ? tiles + blocks : 1) * (long)sizeof({red[4]}));')
            ra2.append(f'p2_{red[3]}_grid')
    b += ['dvmh_amd_mark_kernel_C(lr0, 0);',
          'if (tiles > 0) {',
          IND + 'long wgs = tiles < 16384 ? tiles : 16384;   // persistent: each workgroup walks tiles']
    margs = ['dDesc', 'dStart', 'np', 'tiles', '(int)march', '(int)tune[DVMH_AMD_T_XCD]'] + ra1 + ra2
    for cols, sets in MARCH_NPF.items():
        for v in sets:
            b.append(IND + f'if (mcols == {cols} && npf == {v}) hipLaunchKernelGGL(({P}_pair_kernel<int, {cols}, {v}>), '
                     f'dim3((unsigned)wgs), threads, ldsBytes, stream, {", ".join(margs)});')
    b.append('}')
    bargs = ['dDesc', 'dBlk', 'np', 'blocks', '(int)march', '(int)threads.y', '(int)(threads.x * mcols)'] + ra2
    b += ['if (blocks > 0) {',
          IND + 'long bwgs = blocks < 16384 ? blocks : 16384;',
          IND + f'hipLaunchKernelGGL(({P}_pair_border_kernel<int>), dim3((unsigned)bwgs), dim3(256, 1, 1), 0, stream, {", ".join(bargs)});',
          '}',
          'dvmh_amd_mark_kernel_C(lr0, 1);']
    for ri, red in enumerate(L1.reds):
        loc = f'{red[3]}_grid' if len(red) > 3 else '0'
        b.append(f'dvmh_amd_pair_red_finish_C(batchRef, 0, {ri + 1}, {red[1]}_grid, {loc}, tstart);')
    for ri, red in enumerate(L2.reds):
        loc = f'p2_{red[3]}_grid' if len(red) > 3 else '0'
        b.append(f'dvmh_amd_pair_red_finish_C(batchRef, 1, {ri + 1}, p2_{red[1]}_grid, {loc}, seg2);')
    b.append('free(desc); free(tstart); free(bstart); free(seg2);')
    h.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    h.append('}')
    return '\n'.join(h) + '\n'


def pair_code(L1, L2):
    return '\n'.join([march_kernel(L1, tail=L2), pair_kernels(L1, L2), pair_handler(L1, L2)])


# ----------------------------------------------------------------- GPU: LINE (one dependent axis)
def line_kernel(L, dep):
    """Kernel for a loop whose only dependent loop axis is `dep` (0-based)."""
    r = L.rank
    ind = [a for a in range(r) if a != dep]
    fp, setup = kernel_array_params(L, unit=False)
    for i in range(r):
        fp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    fp += ['IndexT nbx', 'IndexT nby', 'long nblocks', 'int xcd', 'int chunk']
    fp += red_kernel_params(L)
    k = [f'template <typename IndexT>',
         f'__global__ __launch_bounds__(1024) void {L.name}_line{dep}_kernel({", ".join(fp)}) {{']
    b = list(setup)
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    b += red_kernel_locals(L)
    b.append('long tile = dvmh_dev::xcdTile((long)blockIdx.x, nblocks, xcd);')
    b.append('const IndexT bx = (IndexT)(tile % nbx); tile /= nbx;')
    b.append('const IndexT by = (IndexT)(tile % nby); tile /= nby;')
    ax_x = ind[-1] if ind else None
    ax_y = ind[-2] if len(ind) >= 2 else None
    outer = ind[:-2] if len(ind) > 2 else []
    for a in reversed(outer):
        b.append(f'const IndexT t{a} = (IndexT)(tile % cnt{a}); tile /= cnt{a};')
    conds = []
    if ax_x is not None:
        b.append(f'const IndexT t{ax_x} = bx * (IndexT)blockDim.x + (IndexT)threadIdx.x;')
        conds.append(f't{ax_x} < cnt{ax_x}')
    if ax_y is not None:
        b.append(f'const IndexT t{ax_y} = by * (IndexT)blockDim.y + (IndexT)threadIdx.y;')
        conds.append(f't{ax_y} < cnt{ax_y}')
    b.append(f'if ({" && ".join(conds) if conds else "threadIdx.x == 0"}) {{')
    inner = []
    for a in ind:
        inner.append(var_assign(L, a, f't{a}', f'lo{a}', f'st{a}'))
    inner.append(f'for (IndexT t{dep} = 0; t{dep} < cnt{dep}; t{dep}++) {{')
    inner.append(IND + var_assign(L, dep, f't{dep}', f'lo{dep}', f'st{dep}'))
    inner.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 1))
    inner.append('}')
    b += [IND + ln for ln in '\n'.join(inner).split('\n')]
    b.append('}')
    b += red_kernel_tail(L, 'blockIdx.x', accumulate=False)
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    return '\n'.join(k) + '\n'


LINE_REG_CK = [8, 16]


def line_reg_kernel(L, dep):
    """ACROSS with one dependent axis `dep` other than the innermost (threads
    over the independent axes, coalesced along the innermost one): each thread
    walks its line as in line_kernel, but in chunks of CK steps.  The chunk's
    elements of the ACROSS array (and the one ahead) are loaded together into
    a local buffer -- CK + 1 loads in flight instead of one dependent load per
    step -- the body runs against the buffer (RegLineView), and the chunk is
    stored back.  The last element of a full chunk is carried into the next.
    Steps are unrolled, so slot indices are constants and the buffer lives in
    registers.  Loads are clamped into the indices the body reads."""
    r = L.rank
    ind = [a for a in range(r) if a != dep]
    X = L.across['array']
    xct = [p for p in L.arrays if p[1] == X][0][2]
    dmin, dmax = L.stencil_range(X)[dep]
    fp, setup = kernel_array_params(L, unit=False)
    for i in range(r):
        fp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    fp += ['IndexT nbx', 'IndexT nby', 'long nblocks', 'int xcd']
    fp += red_kernel_params(L)
    k = ['template <typename IndexT, int CK, int DIR>',
         f'__global__ __launch_bounds__(256) void {L.name}_lreg{dep}_kernel({", ".join(fp)}) {{']
    b = list(setup)
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    b += red_kernel_locals(L)
    b.append('long tile = dvmh_dev::xcdTile((long)blockIdx.x, nblocks, xcd);')
    b.append('const IndexT bx = (IndexT)(tile % nbx); tile /= nbx;')
    b.append('const IndexT by = (IndexT)(tile % nby); tile /= nby;')
    ax_x = ind[-1]
    ax_y = ind[-2] if len(ind) >= 2 else None
    conds = []
    b.append(f'const IndexT t{ax_x} = bx * (IndexT)blockDim.x + (IndexT)threadIdx.x;')
    conds.append(f't{ax_x} < cnt{ax_x}')
    if ax_y is not None:
        b.append(f'const IndexT t{ax_y} = by * (IndexT)blockDim.y + (IndexT)threadIdx.y;')
        conds.append(f't{ax_y} < cnt{ax_y}')
    conds.append(f'cnt{dep} > 0')   # an empty line reads nothing (the clamp range would be empty)
    b.append(f'if ({" && ".join(conds)}) {{')
    inner = []
    for a in ind:
        inner.append(var_assign(L, a, f't{a}', f'lo{a}', f'st{a}'))
    inner.append(f'{xct} *const lr_row = {X}_base' + ''.join(f' + (IndexT){L.vars[a][0]} * {X}_s{a}' for a in ind) +
                 ';   // this thread\'s line')
    inner.append(f'const IndexT lr_first = lo{dep}, lr_last = lo{dep} + (cnt{dep} - 1) * (IndexT)DIR;')
    inner.append(f'// indices along the line the body reads (offsets {dmin} .. {dmax}): loads are clamped into them')
    inner.append(f'const IndexT lr_rlo = (DIR > 0 ? lr_first : lr_last) + ({dmin}), lr_rhi = (DIR > 0 ? lr_last : lr_first) + ({dmax});')
    inner.append(f'auto lr_ld = [&](IndexT x) -> {xct} {{ x = x < lr_rlo ? lr_rlo : (x > lr_rhi ? lr_rhi : x); return lr_row[x * {X}_s{dep}]; }};')
    inner.append(f'{xct} lr_buf[CK + 2];   // slot 1 + s: step s of the chunk; 0: the element behind; CK + 1: ahead')
    inner.append('lr_buf[0] = lr_ld(lr_first - (IndexT)DIR);')
    inner.append(f'for (IndexT lr_c0 = 0; lr_c0 < cnt{dep}; lr_c0 += CK) {{')
    ch = []
    ch.append(f'const IndexT lr_n = cnt{dep} - lr_c0 < CK ? cnt{dep} - lr_c0 : CK;')
    ch.append(f'const IndexT lr_x0 = lo{dep} + lr_c0 * (IndexT)DIR;   // array index of slot 1')
    ch.append('#pragma unroll')
    ch.append('for (int m = 1; m <= CK + 1; m++) lr_buf[m] = lr_ld(lr_x0 + (IndexT)(m - 1) * (IndexT)DIR);')
    ch.append(f'const dvmh_dev::RegLineView<{xct}, IndexT, {r}, {dep}, DIR> {X}{{lr_buf, lr_x0, 0}};')
    ch.append('#pragma unroll')
    ch.append('for (int lr_s = 0; lr_s < CK; lr_s++) {')
    ch.append(IND + 'if (lr_s < lr_n) {')
    ch.append(IND * 2 + f'const IndexT t{dep} = lr_c0 + lr_s;')
    ch.append(IND * 2 + var_assign(L, dep, f't{dep}', f'lo{dep}', '(IndexT)DIR'))
    ch.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 2))
    ch.append(IND + '}')
    ch.append('}')
    ch.append('#pragma unroll')
    ch.append(f'for (int m = 1; m <= CK; m++) if (m <= lr_n) lr_row[(lr_x0 + (IndexT)(m - 1) * (IndexT)DIR) * {X}_s{dep}] = lr_buf[m];')
    ch.append('lr_buf[0] = lr_buf[CK];   // carried (only a full chunk has a successor)')
    inner += [IND + ln for ln in '\n'.join(ch).split('\n')]
    inner.append('}')
    b += [IND + ln for ln in '\n'.join(inner).split('\n')]
    b.append('}')
    b += red_kernel_tail(L, 'blockIdx.x', accumulate=False)
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    return '\n'.join(k) + '\n'


def line_reg_launch(L, dep):
    r = L.rank
    ind = [a for a in range(r) if a != dep]
    ax_x = ind[-1]
    ax_y = ind[-2] if len(ind) >= 2 else None
    b = [f'DvmType nbx = (cnt{ax_x} + threads.x - 1) / threads.x;',
         f'DvmType nby = {"(cnt%d + threads.y - 1) / threads.y" % ax_y if ax_y is not None else "1"};',
         'DvmType nblocks = nbx * nby;']
    b += handler_reds_prepare(L, 'nblocks', 0)
    b.append('if (nblocks > 0) {')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 0);')
    args = array_fact_args(L, 'int')
    for i in range(r):
        args += [f'(int)boundsLow[{i}]', f'(int)cnt{i}', f'(int)loopSteps[{i}]']
    args += ['(int)nbx', '(int)nby', 'nblocks', '(int)tune[DVMH_AMD_T_XCD]']
    args += red_fact_args(L)
    for ck in LINE_REG_CK:
        for d in (1, -1):
            b.append(IND + f'if (CK == {ck} && loopSteps[{dep}] == {d}) hipLaunchKernelGGL(({L.name}_lreg{dep}_kernel<int, {ck}, {d}>), '
                     f'dim3((unsigned)nblocks), threads, 0, stream, {", ".join(args)});')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 1);')
    b.append('}')
    b += handler_reds_finish(L)
    return b


def line_launch(L, dep):
    r = L.rank
    ind = [a for a in range(r) if a != dep]
    b = []
    ax_x = ind[-1] if ind else None
    ax_y = ind[-2] if len(ind) >= 2 else None
    b.append(f'DvmType nbx = {"(cnt%d + threads.x - 1) / threads.x" % ax_x if ax_x is not None else "1"};')
    b.append(f'DvmType nby = {"(cnt%d + threads.y - 1) / threads.y" % ax_y if ax_y is not None else "1"};')
    nb = 'nbx * nby' + ''.join(f' * cnt{a}' for a in ind[:-2])
    b.append(f'DvmType nblocks = {nb};')
    b += handler_reds_prepare(L, 'nblocks', 0)
    b.append('if (nblocks > 0) {')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 0);')
    for it in ['int', 'long long']:
        args = array_fact_args(L, it)
        for i in range(r):
            args += [f'({it})boundsLow[{i}]', f'({it})cnt{i}', f'({it})loopSteps[{i}]']
        args += [f'({it})nbx', f'({it})nby', 'nblocks', '(int)tune[DVMH_AMD_T_XCD]', '(int)tune[DVMH_AMD_T_CHUNK]']
        args += red_fact_args(L)
        cond = f'kernelIndexT == {"rt_INT" if it == "int" else "rt_LLONG"}'
        b += [IND + s for s in launch_dispatch([(cond, f'{L.name}_line{dep}_kernel<{it}>')], args, 'dim3((unsigned)nblocks)')]
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 1);')
    b.append('}')
    b += handler_reds_finish(L)
    return b


# ----------------------------------------------------------------- GPU: WAVEFRONT (>= 2 dependent axes)
def wave_kernel(L, deps):
    """Hyperplane kernel: iteration-space points with sum of dependent-axis
    positions equal to `plane`; threads over the first (m-1) dependent axes and
    the independent axes; the last dependent position is implied."""
    r = L.rank
    ind = [a for a in range(r) if a not in deps]
    last = deps[-1]
    free = deps[:-1]
    fp, setup = kernel_diag_array_params(L)
    for i in range(r):
        fp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    fp += ['IndexT plane', 'IndexT f0lo', 'IndexT nfx', 'IndexT nfy', 'long nblocks']
    fp += red_kernel_params(L)
    tag = ''.join(map(str, deps))
    k = [f'template <typename IndexT>',
         f'__global__ __launch_bounds__(1024) void {L.name}_wave{tag}_kernel({", ".join(fp)}) {{']
    b = list(setup)
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    b += red_kernel_locals(L)
    # thread x <- the innermost free dependent axis, y <- the next free one, the rest
    # (independent axes, further free axes) from blockIdx.y/z decomposition
    fx = free[-1]
    fy = free[-2] if len(free) >= 2 else None
    rest = (free[:-2] if len(free) > 2 else []) + ind
    b.append(f'long tile = (long)blockIdx.x;')
    b.append(f'const IndexT bx = (IndexT)(tile % nfx); tile /= nfx;')
    b.append(f'const IndexT by = (IndexT)(tile % nfy); tile /= nfy;')
    for a in reversed(rest):
        b.append(f'const IndexT t{a} = (IndexT)(tile % cnt{a}); tile /= cnt{a};')
    b.append(f'const IndexT t{fx} = bx * (IndexT)blockDim.x + (IndexT)threadIdx.x + {"f0lo" if fy is None else "0"};')
    if fy is not None:
        b.append(f'const IndexT t{fy} = by * (IndexT)blockDim.y + (IndexT)threadIdx.y + f0lo;')
    tsum = ' - '.join([f'plane'] + [f't{a}' for a in free])
    b.append(f'const IndexT t{last} = {tsum};')
    conds = [f't{a} < cnt{a}' for a in free] + [f't{last} >= 0', f't{last} < cnt{last}']
    b.append(f'if ({" && ".join(conds)}) {{')
    inner = []
    for i in range(r):
        inner.append(var_assign(L, i, f't{i}', f'lo{i}', f'st{i}'))
    inner.append('{\n' + indent(L.body + '\n', 1) + '}')
    b += [IND + ln for ln in '\n'.join(inner).split('\n')]
    b.append('}')
    b += red_kernel_tail(L, 'blockIdx.x', accumulate=True)
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    return '\n'.join(k) + '\n'


def wave_launch(L, deps):
    r = L.rank
    ind = [a for a in range(r) if a not in deps]
    last = deps[-1]
    free = deps[:-1]
    fx = free[-1]
    fy = free[-2] if len(free) >= 2 else None
    rest = (free[:-2] if len(free) > 2 else []) + ind
    tag = ''.join(map(str, deps))
    b = []
    b.append('DvmType planes = 1' + ''.join(f' + (cnt{a} - 1)' for a in deps) + ';')
    # slot per workgroup of the largest launch: bounding box of the free axes
    b.append(f'DvmType maxFx = (cnt{fx} + threads.x - 1) / threads.x;')
    b.append(f'DvmType maxFy = {"(cnt%d + threads.y - 1) / threads.y" % fy if fy is not None else "1"};')
    b.append('DvmType slots = maxFx * maxFy' + ''.join(f' * cnt{a}' for a in rest) + ';')
    b += handler_reds_prepare(L, 'slots', 1)
    b.append('dvmh_amd_mark_kernel_C(loop_ref, 0);')
    b.append('for (DvmType plane = 0; plane < planes; plane++) {')
    # the first-free-axis range that can meet the plane
    lead = fy if fy is not None else fx
    others = [a for a in deps if a != lead]
    b.append(IND + f'DvmType f0lo = plane - (0' + ''.join(f' + cnt{a} - 1' for a in others) + ');')
    b.append(IND + 'if (f0lo < 0) f0lo = 0;')
    b.append(IND + f'DvmType f0hi = plane < cnt{lead} - 1 ? plane : cnt{lead} - 1;')
    b.append(IND + 'if (f0hi < f0lo) continue;')
    if fy is not None:
        b.append(IND + f'DvmType nfx = (cnt{fx} + threads.x - 1) / threads.x;')
        b.append(IND + 'DvmType nfy = (f0hi - f0lo + 1 + threads.y - 1) / threads.y;')
    else:
        b.append(IND + 'DvmType nfx = (f0hi - f0lo + 1 + threads.x - 1) / threads.x;')
        b.append(IND + 'DvmType nfy = 1;')
    b.append(IND + 'DvmType nblocks = nfx * nfy' + ''.join(f' * cnt{a}' for a in rest) + ';')
    for it in ['int', 'long long']:
        args = array_fact_args(L, it, diag=True)
        for i in range(r):
            args += [f'({it})boundsLow[{i}]', f'({it})cnt{i}', f'({it})loopSteps[{i}]']
        args += [f'({it})plane', f'({it})f0lo', f'({it})nfx', f'({it})nfy', 'nblocks']
        args += red_fact_args(L)
        cond = f'kernelIndexT == {"rt_INT" if it == "int" else "rt_LLONG"}'
        b += [IND + s for s in launch_dispatch([(cond, f'{L.name}_wave{tag}_kernel<{it}>')], args, 'dim3((unsigned)nblocks)')]
    b.append('}')
    b.append('dvmh_amd_mark_kernel_C(loop_ref, 1);')
    b += handler_reds_finish(L)
    return b


def tile_kernel(L):
    """Tiled wavefront for an ACROSS loop whose three axes all depend: tiles of
    TA x TA x TC iterations are run one tile-plane (a + b + c = const) per
    launch.  A workgroup stages its tile of the ACROSS array plus a halo of one
    in LDS (lower halo: final values of earlier tiles; upper halo: values no
    tile has updated yet), runs the tile's local hyperplanes in order with a
    workgroup barrier between them, and writes the tile back.  The body reads
    and writes the ACROSS array through a TileView3 with the usual X[i][j][k]
    syntax; every dependency is met in the sequential order's sense, so the
    results are those of the lexicographic sweep.

    The work on one tile is a device function, called by the per-loop kernel
    (one launch per tile-plane) and by the chain kernel, which runs several
    consecutive sweeps of the loop in one launch sequence: launch t runs
    tile-plane t - LAG * s of sweep s for every sweep s that has one, so a
    sweep stays LAG tile-planes behind the sweep before it."""
    X = L.across['array']
    xct = [p for p in L.arrays if p[1] == X][0][2]
    rng = L.stencil_range(X)   # per axis, the offsets at which the body reads X
    fp, setup = [], []
    for p in L.arrays:
        _, nm, ct, ar = p
        fp.append(f'{ct} *__restrict__ {nm}_base')
        fp += [f'IndexT {nm}_s{k}' for k in range(ar)]
        strides = ', '.join(f'{nm}_s{k}' for k in range(ar))
        setup.append(f'const IndexT {nm}_str[{ar}] = {{{strides}}};')
        vn = f'{nm}_g' if nm == X else nm
        setup.append(f'const auto {vn} = dvmh_dev::makeView<{ct}, IndexT, {ar}, true>({nm}_base, {nm}_str);')
    for p in L.scalars:
        fp.append(f'{p[2]} *__restrict__ {p[1]}_ptr')
        setup.append(scalar_bind(L.body, p))
    for i in range(3):
        fp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    call = []
    for p in L.arrays:
        call += [f'{p[1]}_base'] + [f'{p[1]}_s{q}' for q in range(p[3])]
    for p in L.scalars:
        call.append(f'{p[1]}_ptr')
    for i in range(3):
        call += [f'lo{i}', f'cnt{i}', f'st{i}']
    # ---- the tile: stage, local hyperplanes, write back; reductions by reference
    tp = list(fp) + ['IndexT dt_TA', 'IndexT dt_TC', 'IndexT dt_ta', 'IndexT dt_tb', 'IndexT dt_tc']
    for red in L.reds:
        tp.append(f'{red[2]} &{red[1]}')
        if len(red) > 3:
            tp.append(f'{red[4]} &{red[3]}')
    k = ['template <typename IndexT>',
         f'__device__ static inline void {L.name}_tile_work({", ".join(tp)}) {{']
    b = ['extern __shared__ __align__(16) unsigned char dvmh_lds[];']
    b += setup
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    b += ['const IndexT dt_a0 = dt_ta * dt_TA, dt_b0 = dt_tb * dt_TA, dt_c0 = dt_tc * dt_TC;',
          'const IndexT dt_na = cnt0 - dt_a0 < dt_TA ? cnt0 - dt_a0 : dt_TA;',
          'const IndexT dt_nb = cnt1 - dt_b0 < dt_TA ? cnt1 - dt_b0 : dt_TA;',
          'const IndexT dt_nc = cnt2 - dt_c0 < dt_TC ? cnt2 - dt_c0 : dt_TC;',
          'const IndexT dt_R1 = dt_nb + 2, dt_R2 = dt_nc + 2, dt_R12 = dt_R1 * dt_R2;',
          'const IndexT dt_elems = (dt_na + 2) * dt_R12;',
          f'{xct} *dt_tile = ({xct} *)dvmh_lds;',
          'const IndexT dt_oi = lo0 + dt_a0 - 1, dt_oj = lo1 + dt_b0 - 1, dt_ok = lo2 + dt_c0 - 1;   // array index of tile slot (0, 0, 0)',
          'const IndexT dt_tid = (IndexT)threadIdx.x, dt_nthr = (IndexT)blockDim.x;',
          '// stage the tile and its halo, sixteen loads in flight per thread; an index',
          '// outside those the body reads is clamped into them (its slot is never read)']
    for a, v in enumerate(('dt_ii', 'dt_jj', 'dt_kk')):
        b.append(f'const IndexT {v}_lo = lo{a} + ({rng[a][0]}), {v}_hi = lo{a} + cnt{a} - 1 + ({rng[a][1]});')
    b += ['// slot (p, q, r) of element e = dt_tid + m * dt_nthr, advanced by dt_nthr',
          '// in the radix (R1, R2) of the tile: two divisions per thread, not per element',
          'IndexT dt_sp = dt_nthr / dt_R12, dt_sq = (dt_nthr - dt_sp * dt_R12) / dt_R2, dt_sr = dt_nthr - dt_sp * dt_R12 - dt_sq * dt_R2;',
          'IndexT dt_p = dt_tid / dt_R12, dt_q = (dt_tid - dt_p * dt_R12) / dt_R2, dt_r = dt_tid - dt_p * dt_R12 - dt_q * dt_R2;',
          'for (IndexT dt_b0 = 0; dt_b0 < dt_elems; dt_b0 += 16 * dt_nthr) {',
          IND + 'const IndexT dt_e0 = dt_b0 + dt_tid;',
          IND + f'{xct} dt_v[16];',
          IND + '#pragma unroll',
          IND + 'for (int u = 0; u < 16; u++) {',
          IND * 2 + 'if (dt_b0 + (IndexT)u * dt_nthr >= dt_elems) break;   // (uniform) no element of the workgroup left',
          IND * 2 + '// (past the last element p runs beyond the tile: the clamp keeps the load in bounds)',
          IND * 2 + 'IndexT dt_ii = dt_oi + dt_p, dt_jj = dt_oj + dt_q, dt_kk = dt_ok + dt_r;']
    for v in ('dt_ii', 'dt_jj', 'dt_kk'):
        b.append(IND * 2 + f'{v} = {v} < {v}_lo ? {v}_lo : ({v} > {v}_hi ? {v}_hi : {v});')
    b += [IND * 2 + f'dt_v[u] = {X}_g[dt_ii][dt_jj][dt_kk];',
          IND * 2 + 'dt_r += dt_sr; dt_q += dt_sq; dt_p += dt_sp;',
          IND * 2 + 'if (dt_r >= dt_R2) { dt_r -= dt_R2; dt_q++; }',
          IND * 2 + 'if (dt_q >= dt_R1) { dt_q -= dt_R1; dt_p++; }',
          IND + '}',
          IND + '#pragma unroll',
          IND + 'for (int u = 0; u < 16; u++)',
          IND * 2 + 'if (dt_e0 + (IndexT)u * dt_nthr < dt_elems) dt_tile[dt_e0 + (IndexT)u * dt_nthr] = dt_v[u];',
          '}',
          '__syncthreads();',
          f'const dvmh_dev::TileView3<{xct}, IndexT> {X}{{dt_tile, dt_R1, dt_R2, dt_oi, dt_oj, dt_ok}};',
          '// local hyperplanes p + q + r = s of the tile, in order; a thread\'s (p, q)',
          '// columns do not change from one hyperplane to the next: up to four are',
          '// worked out once (more than four per thread take the generic loop)',
          'const IndexT dt_nab = dt_na * dt_nb;',
          'IndexT dt_myp[4], dt_myq[4];',
          'bool dt_has[4];',
          '#pragma unroll',
          'for (int u = 0; u < 4; u++) {',
          IND + 'const IndexT dt_idx = dt_tid + (IndexT)u * dt_nthr;',
          IND + 'dt_has[u] = dt_idx < dt_nab;',
          IND + 'dt_myp[u] = dt_idx / dt_nb;',
          IND + 'dt_myq[u] = dt_idx - dt_myp[u] * dt_nb;',
          '}',
          'const bool dt_few = dt_nab <= 4 * dt_nthr, dt_one = dt_nab <= dt_nthr;',
          'for (IndexT dt_s = 0; dt_s < dt_na + dt_nb + dt_nc - 2; dt_s++) {',
          IND + 'if (dt_one) {   // at most one column per thread',
          IND * 2 + 'const IndexT dt_p = dt_myp[0], dt_q = dt_myq[0];',
          IND * 2 + 'const IndexT dt_r = dt_s - dt_p - dt_q;',
          IND * 2 + 'if (dt_has[0] && dt_r >= 0 && dt_r < dt_nc) {',
          IND * 3 + 'const IndexT t0 = dt_a0 + dt_p, t1 = dt_b0 + dt_q, t2 = dt_c0 + dt_r;',
          ] + [IND * 3 + var_assign(L, i, f't{i}', f'lo{i}', '(IndexT)1') for i in range(3)] + [indent('{\n' + indent(L.body + '\n', 1) + '}', 3)] + [
          IND * 2 + '}',
          IND + '} else if (dt_few) {',
          IND * 2 + '#pragma unroll',
          IND * 2 + 'for (int u = 0; u < 4; u++) {',
          IND * 3 + 'const IndexT dt_p = dt_myp[u], dt_q = dt_myq[u];',
          IND * 3 + 'const IndexT dt_r = dt_s - dt_p - dt_q;',
          IND * 3 + 'if (dt_has[u] && dt_r >= 0 && dt_r < dt_nc) {',
          IND * 4 + 'const IndexT t0 = dt_a0 + dt_p, t1 = dt_b0 + dt_q, t2 = dt_c0 + dt_r;']
    for i in range(3):
        b.append(IND * 4 + var_assign(L, i, f't{i}', f'lo{i}', '(IndexT)1'))   # the handlers run unit forward steps only
    b.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 4))
    b += [IND * 3 + '}',
          IND * 2 + '}',
          IND + '} else {',
          IND * 2 + 'for (IndexT dt_idx = dt_tid; dt_idx < dt_nab; dt_idx += dt_nthr) {',
          IND * 3 + 'const IndexT dt_p = dt_idx / dt_nb, dt_q = dt_idx - (dt_idx / dt_nb) * dt_nb;',
          IND * 3 + 'const IndexT dt_r = dt_s - dt_p - dt_q;',
          IND * 3 + 'if (dt_r >= 0 && dt_r < dt_nc) {',
          IND * 4 + 'const IndexT t0 = dt_a0 + dt_p, t1 = dt_b0 + dt_q, t2 = dt_c0 + dt_r;']
    for i in range(3):
        b.append(IND * 4 + var_assign(L, i, f't{i}', f'lo{i}', '(IndexT)1'))   # the handlers run unit forward steps only
    b.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 4))
    b += [IND * 3 + '}',
          IND * 2 + '}',
          IND + '}',
          IND + '__syncthreads();',
          '}',
          '// write back the tile (not the halo); (p, q, r) advanced as in the staging',
          'const IndexT dt_nbc = dt_nb * dt_nc;',
          '{',
          IND + 'const IndexT wp = dt_nthr / dt_nbc, wq = (dt_nthr - wp * dt_nbc) / dt_nc, wr = dt_nthr - wp * dt_nbc - wq * dt_nc;',
          IND + 'IndexT p = dt_tid / dt_nbc, q = (dt_tid - p * dt_nbc) / dt_nc, r = dt_tid - p * dt_nbc - q * dt_nc;',
          IND + 'for (IndexT dt_e = dt_tid; dt_e < dt_na * dt_nbc; dt_e += dt_nthr) {',
          IND * 2 + f'{X}_g[dt_oi + 1 + p][dt_oj + 1 + q][dt_ok + 1 + r] = dt_tile[((p + 1) * dt_R1 + (q + 1)) * dt_R2 + (r + 1)];',
          IND * 2 + 'r += wr; q += wq; p += wp;',
          IND * 2 + 'if (r >= dt_nc) { r -= dt_nc; q++; }',
          IND * 2 + 'if (q >= dt_nb) { q -= dt_nb; p++; }',
          IND + '}',
          '}',
          '__syncthreads();   // the LDS tile is free for the next one']
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    redcall = []
    for red in L.reds:
        redcall.append(red[1])
        if len(red) > 3:
            redcall.append(red[3])
    # ---- per-loop kernel: one tile-plane per launch
    kp = list(fp) + ['IndexT dt_TA', 'IndexT dt_TC', 'IndexT dt_ntB', 'IndexT dt_ntC', 'IndexT dt_aLo', 'IndexT dt_tplane']
    kp += red_kernel_params(L)
    g = ['template <typename IndexT>',
         f'__global__ __launch_bounds__(1024) void {L.name}_tile_kernel({", ".join(kp)}) {{']
    gb = ['const IndexT dt_ta = dt_aLo + (IndexT)blockIdx.x / dt_ntB;',
          'const IndexT dt_tb = (IndexT)blockIdx.x - ((IndexT)blockIdx.x / dt_ntB) * dt_ntB;',
          'const IndexT dt_tc = dt_tplane - dt_ta - dt_tb;',
          'if (dt_tc < 0 || dt_tc >= dt_ntC) return;   // not on this tile-plane (uniform per workgroup)']
    gb += red_kernel_locals(L)
    gb.append(f'{L.name}_tile_work<IndexT>({", ".join(call + ["dt_TA", "dt_TC", "dt_ta", "dt_tb", "dt_tc"] + redcall)});')
    gb += red_kernel_tail(L, '((long)(dt_ta * dt_ntB + dt_tb) * dt_ntC + dt_tc)', accumulate=False)
    g.append(indent('\n'.join(gb) + '\n', 1).rstrip('\n'))
    g.append('}')
    out = '\n'.join(k) + '\n\n' + '\n'.join(g) + '\n'
    if L.chain_eligible():
        # ---- chain kernel: sweep s = dt_sLo + blockIdx.y runs tile-plane
        # dt_launch - dt_lag * s; each sweep folds its reductions into its own
        # segment of dt_tiles slots, from its own initial value
        cp = list(fp) + ['IndexT dt_TA', 'IndexT dt_TC', 'IndexT dt_ntA', 'IndexT dt_ntB', 'IndexT dt_ntC',
                         'IndexT dt_launch', 'IndexT dt_sLo', 'IndexT dt_lag', 'long dt_tiles']
        for red in L.reds:
            cp += [f'const {red[2]} *__restrict__ {red[1]}_inits', f'{red[2]} *{red[1]}_grid']
            if len(red) > 3:
                cp += [f'const {red[4]} *__restrict__ {red[3]}_inits', f'{red[4]} *{red[3]}_grid']
        c = ['template <typename IndexT>',
             f'__global__ __launch_bounds__(1024) void {L.name}_chain_kernel({", ".join(cp)}) {{']
        cb = ['const IndexT dt_sw = dt_sLo + (IndexT)blockIdx.y;',
              'const IndexT dt_tplane = dt_launch - dt_lag * dt_sw;',
              'const IndexT dt_aLo0 = dt_tplane - (dt_ntB - 1) - (dt_ntC - 1);',
              'const IndexT dt_aLo = dt_aLo0 > 0 ? dt_aLo0 : 0;',
              'const IndexT dt_aHi = dt_tplane < dt_ntA - 1 ? dt_tplane : dt_ntA - 1;',
              '// workgroups enumerate (a, j) with j over the shorter of the b and c tile',
              '// axes; the other follows from a + b + c = tile-plane',
              'const IndexT dt_nq = dt_ntB < dt_ntC ? dt_ntB : dt_ntC;',
              'const IndexT dt_ta = dt_aLo + (IndexT)blockIdx.x / dt_nq;',
              'const IndexT dt_j = (IndexT)blockIdx.x - ((IndexT)blockIdx.x / dt_nq) * dt_nq;',
              'const IndexT dt_tb = dt_ntB <= dt_ntC ? dt_j : dt_tplane - dt_ta - dt_j;',
              'const IndexT dt_tc = dt_ntB <= dt_ntC ? dt_tplane - dt_ta - dt_j : dt_j;',
              '// this sweep has no tile here (before its first or after its last',
              '// tile-plane, or off the plane): uniform per workgroup',
              'if (dt_tplane < 0 || dt_ta > dt_aHi || dt_tb < 0 || dt_tb >= dt_ntB || dt_tc < 0 || dt_tc >= dt_ntC) return;']
        for red in L.reds:
            cb.append(f'{red[2]} {red[1]} = {red[1]}_inits[dt_sw];')
            if len(red) > 3:
                cb.append(f'{red[4]} {red[3]} = {red[3]}_inits[dt_sw];')
        cb.append(f'{L.name}_tile_work<IndexT>({", ".join(call + ["dt_TA", "dt_TC", "dt_ta", "dt_tb", "dt_tc"] + redcall)});')
        cb += red_kernel_tail(L, '(long)dt_sw * dt_tiles + ((long)(dt_ta * dt_ntB + dt_tb) * dt_ntC + dt_tc)', accumulate=False)
        c.append(indent('\n'.join(cb) + '\n', 1).rstrip('\n'))
        c.append('}')
        out += '\n' + '\n'.join(c) + '\n'
    return out


# Chunk lengths (iterations per staged chunk) the tiled line sweep is
# instantiated for; a thread keeps its share of the next chunk, CK + 1
# elements, in registers while the current one is computed.
LINE_TILE_CK = [8, 16, 32]


def line_tile_kernel(L):
    """Tiled line sweep for an ACROSS loop whose only dependent axis is the
    innermost one.  The direct line kernel gives each thread a line along
    that axis, so neighbouring threads touch addresses a row apart and every
    load is uncoalesced.  Here a workgroup owns blockDim.x lines and walks
    them in chunks of CK iterations, double-buffered in LDS: while each
    thread runs its line's iterations of chunk c in loop order against the
    LDS copy (LineView), the loads of chunk c + 1 are in flight in registers
    (consecutive threads read consecutive elements of a line); then chunk c
    is written back the same coalesced way and chunk c + 1 committed to the
    other buffer.  Slot 0 of a chunk holds the element before it in loop
    order (carried, already updated), slot CK + 1 the element after it (not
    yet updated).  Loads are clamped to the indices the body reads."""
    r = L.rank
    X = L.across['array']
    xct = [p for p in L.arrays if p[1] == X][0][2]
    dlo, dhi = L.stencil_range(X)[r - 1]
    fp, setup = [], []
    for p in L.arrays:
        _, nm, ct, ar = p
        fp.append(f'{ct} *__restrict__ {nm}_base')
        fp += [f'IndexT {nm}_s{k}' for k in range(ar)]
        if nm == X:
            continue
        strides = ', '.join(f'{nm}_s{k}' for k in range(ar))
        setup.append(f'const IndexT {nm}_str[{ar}] = {{{strides}}};')
        setup.append(f'const auto {nm} = dvmh_dev::makeView<{ct}, IndexT, {ar}, false>({nm}_base, {nm}_str);')
    for p in L.scalars:
        fp.append(f'{p[2]} *__restrict__ {p[1]}_ptr')
        setup.append(scalar_bind(L.body, p))
    for i in range(r):
        fp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    fp += ['IndexT dl_nlines']
    fp += red_kernel_params(L)
    kx = r - 1
    # at most 256 threads (the handler checks): lines are few, so registers
    # for deep prefetch matter more than occupancy
    k = ['template <typename IndexT, int CK, int DIR, int REG>',
         f'__global__ __launch_bounds__(256) void {L.name}_ltile_kernel({", ".join(fp)}) {{']
    b = ['extern __shared__ __align__(16) unsigned char dvmh_lds[];']
    b += setup
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    b += red_kernel_locals(L)
    b += ['constexpr IndexT dl_RS = (CK + 2) | 1;   // slots per line row (odd: spreads LDS banks)',
          'const IndexT dl_TL = (IndexT)blockDim.x, dl_tid = (IndexT)threadIdx.x;',
          f'{xct} *dl_buf0 = ({xct} *)dvmh_lds, *dl_buf1 = dl_buf0 + dl_TL * dl_RS;   // two chunks of dl_TL rows',
          'IndexT *dl_rowOff = (IndexT *)(dvmh_lds + (((size_t)2 * dl_TL * dl_RS * sizeof(' + xct + ') + 15) & ~(size_t)15));',
          'const IndexT dl_line0 = (IndexT)blockIdx.x * dl_TL;',
          'const IndexT dl_nl = dl_nlines - dl_line0 < dl_TL ? dl_nlines - dl_line0 : dl_TL;',
          'const bool dl_mine = dl_tid < dl_nl;',
          '// the line this thread owns: its indices on the independent axes']
    if r == 3:
        b += ['const IndexT dl_myLine = dl_line0 + (dl_mine ? dl_tid : 0);',
              'const IndexT t1 = dl_myLine % cnt1, t0 = dl_myLine / cnt1;',
              f'dl_rowOff[dl_tid] = dl_mine ? (lo0 + t0 * st0) * {X}_s0 + (lo1 + t1 * st1) * {X}_s1 : (IndexT)0;']
    else:
        b += ['const IndexT t0 = dl_line0 + (dl_mine ? dl_tid : 0);',
              f'dl_rowOff[dl_tid] = dl_mine ? (lo0 + t0 * st0) * {X}_s0 : (IndexT)0;']
    b += ['__syncthreads();',
          '// indices the body reads along the line: loads are clamped into them',
          f'const IndexT dl_kA = lo{kx}, dl_kZ = lo{kx} + (cnt{kx} - 1) * st{kx};',
          f'const IndexT dl_kMin = (dl_kA < dl_kZ ? dl_kA : dl_kZ) + ({dlo}), dl_kMax = (dl_kA < dl_kZ ? dl_kZ : dl_kA) + ({dhi});',
          f'auto dl_load = [&](IndexT kk, IndexT row) -> {xct} {{',
          IND + f'kk = kk < dl_kMin ? dl_kMin : (kk > dl_kMax ? dl_kMax : kk);',
          IND + f'return {X}_base[dl_rowOff[row] + kk * {X}_s{kx}];',
          '};',
          f'const IndexT dl_nchunks = (cnt{kx} + CK - 1) / CK;',
          '// chunk 0: slots 0 .. n + 1 straight into buffer 0 (none for an empty line: nothing is read)',
          'if (dl_nchunks > 0) {',
          IND + f'const IndexT n = cnt{kx} < CK ? cnt{kx} : CK, w = n + 2;',
          IND + f'const IndexT kb = lo{kx} - st{kx};',
          IND + 'for (IndexT e = dl_tid; e < dl_nl * w; e += dl_TL) {',
          IND * 2 + 'const IndexT row = e / w, sl = e - (e / w) * w;',
          IND * 2 + f'dl_buf0[row * dl_RS + sl] = dl_load(kb + sl * st{kx}, row);',
          IND + '}',
          '}',
          '__syncthreads();',
          f'{xct} dl_pre[CK + 1];   // this thread\'s share of the next chunk (slots 1 .. CK + 1)',
          'for (IndexT dl_c = 0; dl_c < dl_nchunks; dl_c++) {',
          IND + f'{xct} *dl_cur = (dl_c & 1) ? dl_buf1 : dl_buf0, *dl_nxt = (dl_c & 1) ? dl_buf0 : dl_buf1;',
          IND + 'const IndexT dl_c0 = dl_c * CK;',
          IND + f'const IndexT dl_n = cnt{kx} - dl_c0 < CK ? cnt{kx} - dl_c0 : CK;',
          IND + f'const IndexT dl_kb = lo{kx} + (dl_c0 - 1) * st{kx};   // array index held by slot 0',
          IND + '// issue the next chunk\'s loads (slots 1 .. n1 + 1): in flight during the compute',
          IND + f'const IndexT dl_n1 = dl_c + 1 < dl_nchunks ? (cnt{kx} - dl_c0 - CK < CK ? cnt{kx} - dl_c0 - CK : CK) : 0;',
          IND + 'const IndexT dl_w1 = dl_n1 + 1;',
          IND + '// (unconditional: row and index are clamped, so every load is in bounds)',
          IND + '#pragma unroll',
          IND + 'for (int m = 0; m < CK + 1; m++) {',
          IND * 2 + 'const IndexT e = dl_tid + (IndexT)m * dl_TL;',
          IND * 2 + 'const IndexT row0 = e / dl_w1, sl = 1 + (e - row0 * dl_w1);',
          IND * 2 + f'dl_pre[m] = dl_load(dl_kb + (CK + sl) * st{kx}, row0 < dl_nl ? row0 : dl_nl - 1);',
          IND + '}',
          IND + 'if (dl_mine && REG) {',
          IND * 2 + '// the line\'s slots in registers: the recurrence runs at register latency',
          IND * 2 + f'{xct} *dl_line = dl_cur + dl_tid * dl_RS;',
          IND * 2 + f'{xct} dl_reg[CK + 2];',
          IND * 2 + '#pragma unroll',
          IND * 2 + 'for (int m = 0; m < CK + 2; m++) dl_reg[m] = dl_line[m];',
          IND * 2 + f'const IndexT dl_x0 = lo{kx} + dl_c0 * (IndexT)DIR;   // array index of slot 1',
          IND * 2 + f'const dvmh_dev::RegLineView<{xct}, IndexT, {r}, {kx}, DIR> {X}{{dl_reg, dl_x0, 0}};',
          IND * 2 + '#pragma unroll',
          IND * 2 + 'for (int dl_s = 0; dl_s < CK; dl_s++) {',
          IND * 3 + 'if (dl_s < dl_n) {',
          IND * 4 + f'const IndexT t{kx} = dl_c0 + dl_s;']
    for i in range(r):
        b.append(IND * 4 + var_assign(L, i, f't{i}', f'lo{i}', '(IndexT)DIR' if i == kx else f'st{i}'))
    b.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 4))
    b += [IND * 3 + '}',
          IND * 2 + '}',
          IND * 2 + '#pragma unroll',
          IND * 2 + 'for (int m = 1; m <= CK; m++) dl_line[m] = dl_reg[m];',
          IND + '} else if (dl_mine) {',
          IND * 2 + f'const dvmh_dev::LineView<{xct}, IndexT, {r}> {X}{{dl_cur + dl_tid * dl_RS, dl_kb, st{kx}}};',
          IND * 2 + 'for (IndexT dl_i = 1; dl_i <= dl_n; dl_i++) {',
          IND * 3 + f'const IndexT t{kx} = dl_c0 + dl_i - 1;']
    for i in range(r):
        b.append(IND * 3 + var_assign(L, i, f't{i}', f'lo{i}', f'st{i}'))
    b.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 3))
    b += [IND * 2 + '}',
          IND + '}',
          IND + 'if (dl_mine) dl_nxt[dl_tid * dl_RS] = dl_cur[dl_tid * dl_RS + dl_n];   // carried into the next chunk',
          IND + '__syncthreads();',
          IND + '// write back this chunk (slots 1 .. n), commit the next one',
          IND + 'for (IndexT e = dl_tid; e < dl_nl * dl_n; e += dl_TL) {',
          IND * 2 + 'const IndexT row = e / dl_n, sl = 1 + (e - (e / dl_n) * dl_n);',
          IND * 2 + f'{X}_base[dl_rowOff[row] + (dl_kb + sl * st{kx}) * {X}_s{kx}] = dl_cur[row * dl_RS + sl];',
          IND + '}',
          IND + '#pragma unroll',
          IND + 'for (int m = 0; m < CK + 1; m++) {',
          IND * 2 + 'const IndexT e = dl_tid + (IndexT)m * dl_TL;',
          IND * 2 + 'if (dl_n1 > 0 && e < dl_nl * dl_w1) {',
          IND * 3 + 'const IndexT row = e / dl_w1, sl = 1 + (e - (e / dl_w1) * dl_w1);',
          IND * 3 + 'dl_nxt[row * dl_RS + sl] = dl_pre[m];',
          IND * 2 + '}',
          IND + '}',
          IND + '__syncthreads();',
          '}']
    b += red_kernel_tail(L, 'blockIdx.x', accumulate=False)
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    return '\n'.join(k) + '\n'


def line_tile_launch(L):
    r = L.rank
    b = ['DvmType nlines = 1' + ''.join(f' * cnt{a}' for a in range(r - 1)) + ';',
         'DvmType nblocks = (nlines + threads.x - 1) / threads.x;']
    b += handler_reds_prepare(L, 'nblocks', 0)
    b.append('if (nblocks > 0) {')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 0);')
    for it in ['int', 'long long']:
        args = array_fact_args(L, it)
        for i in range(r):
            args += [f'({it})boundsLow[{i}]', f'({it})cnt{i}', f'({it})loopSteps[{i}]']
        args += [f'({it})nlines']
        args += red_fact_args(L)
        for ck in LINE_TILE_CK:
            for d in (1, -1):
                for reg in ((0, 1) if it == 'int' else (0,)):
                    cond = (f'kernelIndexT == {"rt_INT" if it == "int" else "rt_LLONG"} && CK == {ck} && '
                            f'loopSteps[{r - 1}] == {d} && REG == {reg}')
                    b.append(IND + f'if ({cond}) hipLaunchKernelGGL(({L.name}_ltile_kernel<{it}, {ck}, {d}, {reg}>), '
                             f'dim3((unsigned)nblocks), threads, ldsBytes, stream, {", ".join(args)});')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 1);')
    b.append('}')
    b += handler_reds_finish(L)
    return b


def line_vec_kernel(L):
    """Streamed line sweep for an ACROSS loop whose only dependent axis is the
    innermost one, on a unit-stride line: a wave owns 64 lines, a lane walks
    one of them.  Each line is cut into chunks of 128 bytes (CK elements)
    aligned in memory.  A chunk that is whole on every line of the wave moves
    between HBM and registers coalesced -- 8 lanes read one line's chunk as
    16-byte pieces, 8 lines per load -- and is transposed through the wave's
    LDS rows (stride CK + 1), so each lane walks its own line in registers
    (RegLineView: slot 0 the element behind in loop order, already updated;
    CK + 1 the one ahead, not yet updated) and the results go back the same
    way.  The loads of chunk c + 1 are in flight while chunk c is transposed,
    walked and stored (two piece buffers with fixed roles, a loop unrolled by
    two, the first chunk peeled so every wait leaves the previous stores in
    flight).  The chunks before and after that range (partial, or not whole
    on some line of the wave) are walked lane by lane with loads clamped into
    the indices the body reads and stores of the loop's own steps only."""
    r = L.rank
    kx = r - 1
    ind = list(range(r - 1))
    X = L.across['array']
    xct = [p for p in L.arrays if p[1] == X][0][2]
    dmin, dmax = L.stencil_range(X)[kx]
    fp, setup = kernel_array_params(L, unit=False)
    for i in range(r):
        fp += [f'IndexT lo{i}', f'IndexT cnt{i}', f'IndexT st{i}']
    fp += ['IndexT nlines']
    fp += red_kernel_params(L)
    k = ['template <typename IndexT, int DIR>',
         '// a thread per line leaves one or two waves per SIMD: registers go to chunks in flight',
         f'__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void {L.name}_lvec_kernel({", ".join(fp)}) {{']
    b = [f'constexpr int CK = 128 / (int)sizeof({xct});   // elements per chunk: one 128-byte line',
         f'constexpr int LV_W = 16 / (int)sizeof({xct});   // elements per 16-byte piece',
         'constexpr int LV_RS = CK + 1;   // LDS row stride of a line',
         f'typedef {xct} lv_vec __attribute__((ext_vector_type(LV_W)));',
         f'__shared__ __align__(16) {xct} lv_lds[4 * 64 * LV_RS];']
    b += setup
    for v, ct in L.vars:
        b.append(f'{ct} {v};')
    for pr in L.privates:
        b.append(pr + ';')
    b += red_kernel_locals(L)
    b += ['const int lv_lane = (int)(threadIdx.x & 63), lv_sub = lv_lane >> 3, lv_pc = lv_lane & 7;',
          f'{xct} *const lv_buf = lv_lds + (threadIdx.x >> 6) * 64 * LV_RS;   // this wave\'s rows',
          'const IndexT lv_l0 = (IndexT)blockIdx.x * (IndexT)blockDim.x + (IndexT)(threadIdx.x & ~63u);   // the wave\'s first line',
          f'if (lv_l0 < nlines && cnt{kx} > 0) {{   // uniform over the wave']
    inner = ['const IndexT lv_left = nlines - lv_l0;   // lines from the wave\'s first to the last',
             'const bool lv_mine = (IndexT)lv_lane < lv_left;',
             '// row of a line (lines past the end read the last one; nothing of theirs is stored)']
    if r == 3:
        rowof = (f'auto lv_rowOf = [&](IndexT ln) -> {xct} * {{ ln = ln < nlines ? ln : nlines - 1; '
                 f'return {X}_base + (lo0 + (ln / cnt1) * st0) * {X}_s0 + (lo1 + (ln % cnt1) * st1) * {X}_s1; }};')
    else:
        rowof = f'auto lv_rowOf = [&](IndexT ln) -> {xct} * {{ ln = ln < nlines ? ln : nlines - 1; return {X}_base + (lo0 + ln * st0) * {X}_s0; }};'
    inner.append(rowof)
    inner.append('const IndexT lv_line = lv_mine ? lv_l0 + (IndexT)lv_lane : nlines - 1;   // the line this lane walks')
    if r == 3:
        inner.append('const IndexT t1 = lv_line % cnt1, t0 = lv_line / cnt1;')
    else:
        inner.append('const IndexT t0 = lv_line;')
    for a in ind:
        inner.append(var_assign(L, a, f't{a}', f'lo{a}', f'st{a}'))
    inner.append(f'{xct} *const lv_row = lv_rowOf(lv_line);')
    inner += [f'// the body reads the element ahead in loop order (offset {dmax} forward, {dmin} backward)',
              f'constexpr bool LV_AHEAD = DIR > 0 ? ({dmax}) > 0 : ({dmin}) < 0;',
              f'const IndexT lv_kA = lo{kx}, lv_kZ = lo{kx} + (cnt{kx} - 1) * (IndexT)DIR;',
              f'// indices along the line the body reads (offsets {dmin} .. {dmax}): lane loads are clamped into them',
              f'const IndexT lv_rlo = (DIR > 0 ? lv_kA : lv_kZ) + ({dmin}), lv_rhi = (DIR > 0 ? lv_kZ : lv_kA) + ({dmax});',
              f'auto lv_ld = [&](IndexT x) -> {xct} {{ x = x < lv_rlo ? lv_rlo : (x > lv_rhi ? lv_rhi : x); return lv_row[x]; }};',
              f'auto lv_pos = [&]({xct} *row) -> IndexT {{ return (IndexT)(((uintptr_t)(row + lv_kA) / sizeof({xct})) & (uintptr_t)(CK - 1)); }};',
              '// lv_v loop steps precede the first one in its aligned chunk; chunk c holds',
              '// steps c * CK - lv_v .. + CK - 1 in slots 1 .. CK; chunks up to lv_cL are whole from chunk 1 on',
              'const IndexT lv_p = lv_pos(lv_row);',
              'const IndexT lv_v = DIR > 0 ? lv_p : (IndexT)(CK - 1) - lv_p;',
              f'const IndexT lv_nch = (lv_v + cnt{kx} + (IndexT)(CK - 1)) / (IndexT)CK;',
              f'const IndexT lv_cL = (lv_v + cnt{kx}) / (IndexT)CK - 1;',
              'auto lv_x0 = [&](IndexT lv_cc) -> IndexT { return lv_kA + (lv_cc * (IndexT)CK - lv_v) * (IndexT)DIR; };   // index of slot 1',
              '// chunks 1 .. lv_w1 are whole on every line of the wave',
              'IndexT lv_w1 = lv_mine ? lv_cL : (IndexT)0x7fffffff;',
              'for (int o = 32; o > 0; o >>= 1) { const IndexT q = __shfl_xor(lv_w1, o); lv_w1 = q < lv_w1 ? q : lv_w1; }',
              'lv_w1 = __builtin_amdgcn_readfirstlane(lv_w1);',
              f'{xct} lv_carry = lv_ld(lv_x0(0) - (IndexT)DIR), lv_ah = 0;']

    def walk(c, guard):
        # the loop variable is formed from the chunk's slot-1 index, the view's
        # base: the slot arithmetic then folds to a constant in every step
        s = ['const IndexT lv_xs = lv_x0(' + c + ');',
             f'const dvmh_dev::RegLineView<{xct}, IndexT, {r}, {kx}, DIR> {X}{{lv_w, lv_xs, 0}};',
             '#pragma unroll',
             'for (int lv_s = 0; lv_s < CK; lv_s++) {']
        if guard:
            s.append(IND + f'const IndexT t{kx} = ({c}) * (IndexT)CK - lv_v + (IndexT)lv_s;')
            s.append(IND + f'if (t{kx} >= 0 && t{kx} < cnt{kx}) {{')
        d = 2 if guard else 1
        s.append(IND * d + var_assign(L, kx, '(IndexT)lv_s', 'lv_xs', '(IndexT)DIR'))
        s.append(indent('{\n' + indent(L.body + '\n', 1) + '}', d))
        if guard:
            s.append(IND + '}')
        s.append('}')
        return s

    def partial(c):
        s = ['{',
             IND + f'const IndexT x0 = lv_x0({c});',
             IND + f'{xct} lv_w[CK + 2];',
             IND + 'lv_w[0] = lv_carry;',
             IND + '#pragma unroll',
             IND + 'for (int m = 0; m < CK + 1; m++) lv_w[1 + m] = lv_ld(x0 + (IndexT)m * (IndexT)DIR);']
        s += [IND + ln for ln in walk(c, True)]
        s += [IND + '#pragma unroll',
              IND + 'for (int m = 0; m < CK; m++) {',
              IND * 2 + f'const IndexT t = ({c}) * (IndexT)CK - lv_v + (IndexT)m;',
              IND * 2 + f'if (t >= 0 && t < cnt{kx}) lv_row[x0 + (IndexT)m * (IndexT)DIR] = lv_w[1 + m];',
              IND + '}',
              IND + 'lv_carry = lv_w[CK];',
              '}']
        return s

    sync = '__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");'

    def step(c, cur, nxt, full):
        # chunk c (pieces in `cur`) through LDS and the walk; chunk c + 1 into
        # `nxt`.  full: every line of the wave exists, so the walk and the
        # stores need no guard (a guarded store may be skipped, and the wait
        # for the next chunk's loads would then also wait for the stores)
        walk_lines = walk('lv_c', False)
        if full:
            walk_lines = [IND + ln for ln in walk_lines]
        else:
            walk_lines = [IND + 'if (lv_mine) {'] + [IND * 2 + ln for ln in walk_lines] + [IND + '}']
        store = '*(lv_vec *)(lv_pb[m] + (IndexT)DIR * lv_c * (IndexT)CK + lv_pc * LV_W) = u;'
        if not full:
            store = 'if ((IndexT)(m * 8 + lv_sub) < lv_left) ' + store
        return (['{',
                 IND + f'const IndexT lv_c = {c};',
                 IND + f'lv_gload(lv_c + 1 <= lv_w1 ? lv_c + 1 : lv_w1, {nxt});   // the last chunk reloads itself: no branch',
                 IND + 'if constexpr (LV_AHEAD) lv_ah = lv_ld(lv_x0(lv_c) + (IndexT)CK * (IndexT)DIR);',
                 IND + '__builtin_amdgcn_sched_barrier(0);   // those loads issue before this chunk\'s wait',
                 IND + '#pragma unroll',
                 IND + 'for (int m = 0; m < 8; m++)',
                 IND * 2 + '#pragma unroll',
                 IND * 2 + f'for (int e = 0; e < LV_W; e++) lv_buf[(m * 8 + lv_sub) * LV_RS + lv_pc * LV_W + e] = {cur}[m][e];',
                 IND + sync,
                 IND + f'{xct} lv_w[CK + 2];',
                 IND + 'lv_w[0] = lv_carry;',
                 IND + '#pragma unroll',
                 IND + 'for (int q = 0; q < CK; q++) lv_w[1 + q] = lv_buf[lv_lane * LV_RS + (DIR > 0 ? q : CK - 1 - q)];',
                 IND + 'lv_w[CK + 1] = lv_ah;'] +
                walk_lines + [
                 IND + 'lv_carry = lv_w[CK];',
                 IND + '#pragma unroll',
                 IND + 'for (int q = 0; q < CK; q++) lv_buf[lv_lane * LV_RS + (DIR > 0 ? q : CK - 1 - q)] = lv_w[1 + q];',
                 IND + sync,
                 IND + '#pragma unroll',
                 IND + 'for (int m = 0; m < 8; m++) {',
                 IND * 2 + 'lv_vec u;',
                 IND * 2 + '#pragma unroll',
                 IND * 2 + 'for (int e = 0; e < LV_W; e++) u[e] = lv_buf[(m * 8 + lv_sub) * LV_RS + lv_pc * LV_W + e];',
                 IND * 2 + store,
                 IND + '}',
                 IND + sync,
                 '}'])

    def head_step(full):
        # chunk 0 through LDS like a whole chunk, its walk guarded to the
        # loop's steps; elements that are no step go back unchanged
        walk_lines = walk('(IndexT)0', True)
        if full:
            walk_lines = [IND + ln for ln in walk_lines]
        else:
            walk_lines = [IND + 'if (lv_mine) {'] + [IND * 2 + ln for ln in walk_lines] + [IND + '}']
        store = '*(lv_vec *)(lv_pb[m] + lv_pc * LV_W) = u;'
        if not full:
            store = 'if ((IndexT)(m * 8 + lv_sub) < lv_left) ' + store
        return (['{',
                 IND + 'lv_vec lv_f[8];',
                 IND + 'lv_gload(0, lv_f);',
                 IND + 'if constexpr (LV_AHEAD) lv_ah = lv_ld(lv_x0(0) + (IndexT)CK * (IndexT)DIR);',
                 IND + '#pragma unroll',
                 IND + 'for (int m = 0; m < 8; m++)',
                 IND * 2 + '#pragma unroll',
                 IND * 2 + 'for (int e = 0; e < LV_W; e++) lv_buf[(m * 8 + lv_sub) * LV_RS + lv_pc * LV_W + e] = lv_f[m][e];',
                 IND + sync,
                 IND + f'{xct} lv_w[CK + 2];',
                 IND + 'lv_w[0] = lv_carry;',
                 IND + '#pragma unroll',
                 IND + 'for (int q = 0; q < CK; q++) lv_w[1 + q] = lv_buf[lv_lane * LV_RS + (DIR > 0 ? q : CK - 1 - q)];',
                 IND + 'lv_w[CK + 1] = lv_ah;'] +
                walk_lines + [
                 IND + 'lv_carry = lv_w[CK];',
                 IND + '#pragma unroll',
                 IND + 'for (int q = 0; q < CK; q++) lv_buf[lv_lane * LV_RS + (DIR > 0 ? q : CK - 1 - q)] = lv_w[1 + q];',
                 IND + sync,
                 IND + '#pragma unroll',
                 IND + 'for (int m = 0; m < 8; m++) {',
                 IND * 2 + 'lv_vec u;',
                 IND * 2 + '#pragma unroll',
                 IND * 2 + 'for (int e = 0; e < LV_W; e++) u[e] = lv_buf[(m * 8 + lv_sub) * LV_RS + lv_pc * LV_W + e];',
                 IND * 2 + store,
                 IND + '}',
                 IND + sync,
                 '}'])

    def pipeline(full):
        p = ['lv_vec lv_g[8], lv_h[8];',
             'lv_gload(1, lv_h);']
        p += step('(IndexT)1', 'lv_h', 'lv_g', full)
        p += ['IndexT lv_c2 = 2;   // lv_g holds chunk lv_c2 while lv_c2 <= lv_w1',
              'while (lv_c2 + 1 <= lv_w1) {']
        p += [IND + ln for ln in step('lv_c2', 'lv_g', 'lv_h', full)]
        p += [IND + ln for ln in step('lv_c2 + 1', 'lv_h', 'lv_g', full)]
        p += [IND + 'lv_c2 += 2;', '}', 'if (lv_c2 == lv_w1)']
        p += [IND + ln for ln in step('lv_c2', 'lv_g', 'lv_h', full)]
        return p

    inner.append('if (lv_w1 >= 1) {')
    pipe = ['// this lane\'s pieces: line lv_l0 + m * 8 + lv_sub, 16-byte piece lv_pc of every chunk;',
            '// lv_pb[m] is that line\'s chunk 0 (aligned), chunk c at lv_pb[m] + DIR * c * CK',
            f'{xct} *lv_pb[8];',
            '#pragma unroll',
            'for (int m = 0; m < 8; m++) {',
            IND + f'{xct} *const row = lv_rowOf((IndexT)(m * 8 + lv_sub) < lv_left ? lv_l0 + (IndexT)(m * 8 + lv_sub) : nlines - 1);',
            IND + 'lv_pb[m] = row + lv_kA - lv_pos(row);',
            '}',
            'auto lv_gload = [&](IndexT c, lv_vec (&g)[8]) {',
            IND + '#pragma unroll',
            IND + 'for (int m = 0; m < 8; m++) g[m] = *(const lv_vec *)(lv_pb[m] + (IndexT)DIR * c * (IndexT)CK + lv_pc * LV_W);',
            '};',
            '// the first chunk is partial somewhere (or wholly before the loop\'s first step).  When it',
            '// lies inside the indices the body reads on every line of the wave, it goes through LDS',
            '// too (a walk guarded to the loop\'s steps); otherwise lane by lane',
            'const IndexT lv_lo0 = DIR > 0 ? lv_x0(0) : lv_x0(0) - (IndexT)(CK - 1);',
            'if (__all(!lv_mine || (lv_lo0 >= lv_rlo && lv_lo0 + (IndexT)(CK - 1) <= lv_rhi))) {',
            IND + 'if (lv_left >= 64) {']
    pipe += [IND * 2 + ln for ln in head_step(True)]
    pipe += [IND + '} else {']
    pipe += [IND * 2 + ln for ln in head_step(False)]
    pipe += [IND + '}',
             '} else if (lv_mine)']
    pipe += [IND + ln for ln in partial('(IndexT)0')]
    pipe.append('if (lv_left >= 64) {   // every line of the wave exists')
    pipe += [IND + ln for ln in pipeline(True)]
    pipe.append('} else {')
    pipe += [IND + ln for ln in pipeline(False)]
    pipe.append('}')
    inner += [IND + ln for ln in pipe]
    inner.append('}')
    inner.append('// chunks not whole on every line of the wave: lane by lane')
    inner.append('if (lv_mine)')
    inner.append(IND + 'for (IndexT lv_c = lv_w1 >= 1 ? lv_w1 + 1 : 0; lv_c < lv_nch; lv_c++)')
    inner += [IND * 2 + ln for ln in partial('lv_c')]
    b += [IND + ln for ln in '\n'.join(inner).split('\n')]
    b.append('}')
    b += red_kernel_tail(L, 'blockIdx.x', accumulate=False)
    k.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    k.append('}')
    return '\n'.join(k) + '\n'


def line_vec_launch(L):
    r = L.rank
    b = ['DvmType nlines = 1' + ''.join(f' * cnt{a}' for a in range(r - 1)) + ';',
         'DvmType nblocks = (nlines + threads.x - 1) / threads.x;']
    b += handler_reds_prepare(L, 'nblocks', 0)
    b.append('if (nblocks > 0) {')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 0);')
    args = array_fact_args(L, 'int')
    for i in range(r):
        args += [f'(int)boundsLow[{i}]', f'(int)cnt{i}', f'(int)loopSteps[{i}]']
    args += ['(int)nlines']
    args += red_fact_args(L)
    for d in (1, -1):
        b.append(IND + f'if (loopSteps[{r - 1}] == {d}) hipLaunchKernelGGL(({L.name}_lvec_kernel<int, {d}>), '
                 f'dim3((unsigned)nblocks), threads, 0, stream, {", ".join(args)});')
    b.append(IND + 'dvmh_amd_mark_kernel_C(loop_ref, 1);')
    b.append('}')
    b += handler_reds_finish(L)
    return b


def tile_launch(L):
    X = L.across['array']
    xct = [p for p in L.arrays if p[1] == X][0][2]
    b = ['DvmType TA = tune[DVMH_AMD_T_TILE_A], TC = tune[DVMH_AMD_T_TILE_B];',
         'DvmType ntA = (cnt0 + TA - 1) / TA, ntB = (cnt1 + TA - 1) / TA, ntC = (cnt2 + TC - 1) / TC;',
         f'size_t ldsBytes = (size_t)(TA + 2) * (TA + 2) * (TC + 2) * sizeof({xct});',
         'DvmType tiles = ntA * ntB * ntC;']
    b += handler_reds_prepare(L, 'tiles', 0)
    b.append('dvmh_amd_mark_kernel_C(loop_ref, 0);')
    b.append('for (DvmType tplane = 0; tplane <= (ntA - 1) + (ntB - 1) + (ntC - 1); tplane++) {')
    b.append(IND + 'DvmType aLo = tplane - (ntB - 1) - (ntC - 1);')
    b.append(IND + 'if (aLo < 0) aLo = 0;')
    b.append(IND + 'DvmType aHi = tplane < ntA - 1 ? tplane : ntA - 1;')
    b.append(IND + 'DvmType grid = (aHi - aLo + 1) * ntB;')
    for it in ['int', 'long long']:
        args = array_fact_args(L, it)
        for i in range(3):
            args += [f'({it})boundsLow[{i}]', f'({it})cnt{i}', f'({it})loopSteps[{i}]']
        args += [f'({it})TA', f'({it})TC', f'({it})ntB', f'({it})ntC', f'({it})aLo', f'({it})tplane']
        args += red_fact_args(L)
        cond = f'kernelIndexT == {"rt_INT" if it == "int" else "rt_LLONG"}'
        b.append(IND + f'if ({cond}) hipLaunchKernelGGL(({L.name}_tile_kernel<{it}>), dim3((unsigned)grid), threads, ldsBytes, '
                 f'stream, {", ".join(args)});')
    b.append('}')
    b.append('dvmh_amd_mark_kernel_C(loop_ref, 1);')
    b += handler_reds_finish(L)
    return b


# Tile-planes by which a sweep of a chain trails the sweep before it: tile
# (a, b, c) reads only the tiles one tile-plane below and above it (every
# access of the ACROSS array moves along one axis by at most one), so sweep s
# may run tile-plane p once sweep s - 1 has finished tile-plane p + 1; with a
# lag of two, the tiles two sweeps run in one launch are never neighbours.
CHAIN_LAG = 2


def chain_handler(L):
    """Chain handler of a tiled-wavefront ACROSS loop: the group of
    consecutive sweeps dvmh_amd_batch_end_C hands over runs as one launch
    sequence, sweep s CHAIN_LAG tile-planes behind sweep s - 1, each sweep with
    its own reduction segment.  Sweeps that differ in bounds, arrays, scalars
    or layout, or a configuration other than the tiled wavefront, run one
    after another through the ordinary handler."""
    X = L.across['array']
    xct = [p for p in L.arrays if p[1] == X][0][2]
    h = [f'extern "C" void {L.name}_cuda_chain(DvmType *pBatchRef) {{']
    b = ['DvmType batchRef = *pBatchRef;',
         'long nl = (long)dvmh_amd_batch_count_C(batchRef);',
         'DvmType loop_ref = dvmh_amd_batch_loop_C(batchRef, 0);',
         'DvmType tune[DVMH_AMD_TUNE_N];',
         '// the chain has its own tuner site: tile shapes timed per sweep of a chain',
         'dvmh_amd_loop_get_tuning_C(loop_ref, DVMH_AMD_KIND_CHAIN, tune);',
         'dim3 threads(0, 0, 0);',
         'hipStream_t stream;',
         *regs_query(f'{L.name}_chain_kernel<int>'),
         'DvmType dependency_mask = dvmh_loop_get_dependency_mask_C(loop_ref);',
         'int kernelIndexT = (int)dvmh_loop_guess_index_type_C(loop_ref);',
         'if (kernelIndexT == rt_LONG) kernelIndexT = rt_LLONG;',
         'DvmType boundsLow[3], boundsHigh[3], loopSteps[3];',
         'dvmh_loop_fill_bounds_C(loop_ref, boundsLow, boundsHigh, loopSteps);',
         'DvmType TA = tune[DVMH_AMD_T_TILE_A], TC = tune[DVMH_AMD_T_TILE_B];',
         f'size_t ldsBytes = (size_t)(TA + 2) * (TA + 2) * (TC + 2) * sizeof({xct});',
         'bool chained = tune[DVMH_AMD_T_VARIANT] == 1 && dependency_mask == 7 && TA >= 2 && TC >= 2 && ldsBytes <= 64512 && '
         'threads.y == 1 && threads.z == 1 && loopSteps[0] == 1 && loopSteps[1] == 1 && loopSteps[2] == 1;']
    for pi, p in enumerate(L.params):
        if p[0] == 'array':
            _, nm, ct, ar = p
            b.append(f'DvmType *{nm}_hdr = (DvmType *)dvmh_amd_batch_param_C(batchRef, 0, {pi});')
            b.append(f'{ct} *{nm} = ({ct} *)dvmh_get_natural_base_C(1, {nm}_hdr);')
            b.append(f'DvmType {nm}_devHdr[64 + 7];')
            b.append(f'if (dvmh_fill_header_C(1, {nm}, {nm}_hdr, {nm}_devHdr, 0) != 0) chained = false;')
        else:
            b.append(f'{p[2]} *{p[1]} = ({p[2]} *)dvmh_get_device_addr_C(1, dvmh_amd_batch_param_C(batchRef, 0, {pi}));')
    # every sweep of the chain: same arrays, scalars and bounds as the first
    b.append('for (long li = 1; li < nl && chained; li++) {')
    b.append(IND + 'DvmType lr = dvmh_amd_batch_loop_C(batchRef, li);')
    b.append(IND + 'DvmType lo[3], hi[3], st[3];')
    b.append(IND + 'dvmh_loop_fill_bounds_C(lr, lo, hi, st);')
    b.append(IND + 'for (int a = 0; a < 3; a++) if (lo[a] != boundsLow[a] || hi[a] != boundsHigh[a] || st[a] != loopSteps[a]) chained = false;')
    for pi, p in enumerate(L.params):
        b.append(IND + f'if (dvmh_amd_batch_param_C(batchRef, li, {pi}) != dvmh_amd_batch_param_C(batchRef, 0, {pi})) chained = false;')
    b.append('}')
    b.append('if (!chained) {')
    b.append(IND + '// one sweep after another through the ordinary handler')
    b.append(IND + 'for (long li = 0; li < nl; li++) {')
    b.append(IND * 2 + 'DvmType lr = dvmh_amd_batch_loop_C(batchRef, li);')
    args = ['&lr']
    for pi, p in enumerate(L.params):
        cast = 'DvmType *' if p[0] == 'array' else f'{p[2]} *'
        args.append(f'({cast})dvmh_amd_batch_param_C(batchRef, li, {pi})')
    b.append(IND * 2 + f'{L.name}_cuda({", ".join(args)});')
    b.append(IND + '}')
    b.append(IND + 'return;')
    b.append('}')
    b += ['DvmType cnt0 = boundsHigh[0] - boundsLow[0] + 1, cnt1 = boundsHigh[1] - boundsLow[1] + 1, cnt2 = boundsHigh[2] - boundsLow[2] + 1;',
          'DvmType ntA = (cnt0 + TA - 1) / TA, ntB = (cnt1 + TA - 1) / TA, ntC = (cnt2 + TC - 1) / TC;',
          'DvmType tiles = ntA * ntB * ntC, planes = ntA + ntB + ntC - 2;',
          'DvmType *segStart = (DvmType *)malloc((size_t)(nl + 1) * sizeof(DvmType));',
          'for (long li = 0; li <= nl; li++) segStart[li] = li * tiles;']
    ra = []
    for ri, red in enumerate(L.reds):
        ct = red[2]
        b.append(f'{ct} *{red[1]}_grid = ({ct} *)dvmh_amd_batch_scratch_C(batchRef, nl * tiles * (long)sizeof({ct}));')
        b.append(f'{ct} *{red[1]}_inits = ({ct} *)dvmh_amd_batch_scratch_C(batchRef, nl * (long)sizeof({ct}));')
        b.append(f'{ct} *{red[1]}_ih = ({ct} *)malloc((size_t)nl * sizeof({ct}));')
        if len(red) > 3:
            lt = red[4]
            b.append(f'{lt} *{red[3]}_grid = ({lt} *)dvmh_amd_batch_scratch_C(batchRef, nl * tiles * (long)sizeof({lt}));')
            b.append(f'{lt} *{red[3]}_inits = ({lt} *)dvmh_amd_batch_scratch_C(batchRef, nl * (long)sizeof({lt}));')
            b.append(f'{lt} *{red[3]}_ih = ({lt} *)malloc((size_t)nl * sizeof({lt}));')
            b.append(f'for (long li = 0; li < nl; li++) dvmh_loop_red_init_C(dvmh_amd_batch_loop_C(batchRef, li), {ri + 1}, &{red[1]}_ih[li], &{red[3]}_ih[li]);')
            b.append(f'dvmh_amd_batch_upload_C(batchRef, {red[3]}_inits, {red[3]}_ih, nl * (long)sizeof({lt}));')
        else:
            b.append(f'for (long li = 0; li < nl; li++) dvmh_loop_red_init_C(dvmh_amd_batch_loop_C(batchRef, li), {ri + 1}, &{red[1]}_ih[li], 0);')
        b.append(f'dvmh_amd_batch_upload_C(batchRef, {red[1]}_inits, {red[1]}_ih, nl * (long)sizeof({ct}));')
        ra += [f'{red[1]}_inits', f'{red[1]}_grid']
        if len(red) > 3:
            ra += [f'{red[3]}_inits', f'{red[3]}_grid']
    b += [f'const DvmType lag = {CHAIN_LAG};',
          'dvmh_amd_mark_kernel_C(loop_ref, 0);',
          'for (DvmType t = 0; t < planes + lag * (nl - 1); t++) {',
          IND + '// sweeps with a tile-plane in this launch: 0 <= t - lag * s < planes',
          IND + 'DvmType sLo = t - (planes - 1) <= 0 ? 0 : (t - (planes - 1) + lag - 1) / lag;',
          IND + 'DvmType sHi = t / lag < nl - 1 ? t / lag : nl - 1;',
          IND + 'if (sLo > sHi) continue;',
          IND + 'DvmType rows = 0;',
          IND + 'for (DvmType sw = sLo; sw <= sHi; sw++) {',
          IND * 2 + 'DvmType p = t - lag * sw, aLo = p - (ntB - 1) - (ntC - 1), aHi = p < ntA - 1 ? p : ntA - 1;',
          IND * 2 + 'if (aLo < 0) aLo = 0;',
          IND * 2 + 'if (aHi - aLo + 1 > rows) rows = aHi - aLo + 1;',
          IND + '}',
          IND + 'dim3 grid((unsigned)(rows * (ntB < ntC ? ntB : ntC)), (unsigned)(sHi - sLo + 1), 1);']
    for it in ['int', 'long long']:
        args = array_fact_args(L, it)
        for i in range(3):
            args += [f'({it})boundsLow[{i}]', f'({it})cnt{i}', f'({it})loopSteps[{i}]']
        args += [f'({it})TA', f'({it})TC', f'({it})ntA', f'({it})ntB', f'({it})ntC', f'({it})t', f'({it})sLo', f'({it})lag',
                 '(long)tiles'] + ra
        cond = f'kernelIndexT == {"rt_INT" if it == "int" else "rt_LLONG"}'
        b.append(IND + f'if ({cond}) hipLaunchKernelGGL(({L.name}_chain_kernel<{it}>), grid, threads, ldsBytes, stream, {", ".join(args)});')
    b.append('}')
    b.append('dvmh_amd_mark_kernel_C(loop_ref, 1);')
    for ri, red in enumerate(L.reds):
        loc = f'{red[3]}_grid' if len(red) > 3 else '0'
        b.append(f'dvmh_amd_batch_red_finish_C(batchRef, {ri + 1}, {red[1]}_grid, {loc}, segStart);')
        b.append(f'free({red[1]}_ih);')
        if len(red) > 3:
            b.append(f'free({red[3]}_ih);')
    b.append('free(segStart);')
    h.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    h.append('}')
    return '\n'.join(h) + '\n'


def across_handler(L):
    r = L.rank
    ndep = len(L.dep_axes())
    h = [f'extern "C" void {L.name}_cuda(DvmType *pLoopRef{"".join(", DvmType " + p[1] + "_hdr[]" if p[0] == "array" else ", " + p[2] + " *" + p[1] + "_ptr" for p in L.params)}) {{']
    b = ['DvmType dependency_mask = dvmh_loop_get_dependency_mask_C(*pLoopRef);',
         'int dependency_num = __builtin_popcountl((unsigned long)dependency_mask);',
         f'DvmType kind = dependency_num <= 1 ? DVMH_AMD_KIND_LINE : DVMH_AMD_KIND_WAVEFRONT;']
    pro = handler_prologue(L, 'kind', want_ext=True, autotfm=True, regs_kernel=f'{L.name}_line0_kernel<int>')
    b += pro
    for p in L.arrays:
        _, nm, ct, ar = p
        for it in ['int', 'long long']:
            t = it.replace(' ', '_')
            b.append(f'dvmh_dev::DiagInfo<{it}> {nm}_di_{t};')
            b.append(f'{nm}_di_{t}.xAxis = {nm}_tfm == 2 ? (int){nm}_devHdr[64 + 0] : 0;')
            b.append(f'{nm}_di_{t}.yAxis = {nm}_tfm == 2 ? (int){nm}_devHdr[64 + 3] : 0;')
            b.append(f'{nm}_di_{t}.slash = {nm}_tfm == 2 ? (int){nm}_devHdr[64 + 6] : 0;')
            b.append(f'{nm}_di_{t}.xFirst = ({it})({nm}_tfm == 2 ? {nm}_devHdr[64 + 1] : 0);')
            b.append(f'{nm}_di_{t}.xLen = ({it})({nm}_tfm == 2 ? {nm}_devHdr[64 + 2] : 1);')
            b.append(f'{nm}_di_{t}.yFirst = ({it})({nm}_tfm == 2 ? {nm}_devHdr[64 + 4] : 0);')
            b.append(f'{nm}_di_{t}.yLen = ({it})({nm}_tfm == 2 ? {nm}_devHdr[64 + 5] : 1);')
    for i in range(r):
        b.append(f'DvmType cnt{i} = (boundsHigh[{i}] - boundsLow[{i}]) / loopSteps[{i}] + 1;')
    first = True
    # one dependent axis: line kernels for each possible axis
    for dep in range(r):
        mask = 1 << (r - 1 - dep)
        kw = 'if' if first else '} else if'
        first = False
        b.append(f'{kw} (dependency_mask == {mask}) {{')
        for p in L.arrays:
            b.append(IND + f'if ({p[1]}_tfm == 2) {{ fprintf(stderr, "{L.name}: diagonal layout in a line sweep\\n"); abort(); }}')
        if dep == r - 1 and L.line_tile_eligible():
            X = L.across['array']
            xct = [p for p in L.arrays if p[1] == X][0][2]
            b.append(IND + '// tiled line sweep: unit steps along the line, natural layout, chunk fits 64 KB of LDS')
            b.append(IND + '// REG (TILE_B = 1): the walk on a register copy, 32-bit indices only (the slot arithmetic must fold)')
            b.append(IND + 'DvmType CK = tune[DVMH_AMD_T_TILE_A], RS = (CK + 2) | 1, REG = tune[DVMH_AMD_T_TILE_B] && kernelIndexT == rt_INT ? 1 : 0;')
            b.append(IND + f'size_t ldsBytes = (((size_t)2 * threads.x * RS * sizeof({xct}) + 15) & ~(size_t)15) + (size_t)threads.x * sizeof(DvmType);')
            cks = ' || '.join(f'CK == {ck}' for ck in LINE_TILE_CK)
            b.append(IND + f'bool ltile = tune[DVMH_AMD_T_VARIANT] == 1 && (loopSteps[{r - 1}] == 1 || loopSteps[{r - 1}] == -1) && '
                     f'{X}_tfm == 0 && threads.x <= 256 && threads.y == 1 && threads.z == 1 && ({cks}) && ldsBytes <= 64512;')
            b.append(IND + '// streamed line sweep (variant 3): unit stride along the line, 32-bit indices, whole waves, no dynamic LDS')
            b.append(IND + f'bool lvec = tune[DVMH_AMD_T_VARIANT] == 3 && (loopSteps[{r - 1}] == 1 || loopSteps[{r - 1}] == -1) && '
                     f'{X}_tfm == 0 && {X}_devHdr[{r}] == 1 && kernelIndexT == rt_INT && threads.x <= 256 && threads.x % 64 == 0 && '
                     f'threads.y == 1 && threads.z == 1 && sizeof({xct}) >= 4 && sizeof({xct}) <= 8;')
            b.append(IND + 'if (lvec) {')
            b += [IND * 2 + s for s in line_vec_launch(L)]
            b.append(IND + '} else if (ltile) {')
            b += [IND * 2 + s for s in line_tile_launch(L)]
            b.append(IND + '} else {')
            b += [IND * 2 + s for s in line_launch(L, dep)]
            b.append(IND + '}')
        elif dep < r - 1 and L.line_stage_eligible(dep):
            X = L.across['array']
            cks = ' || '.join(f'CK == {ck}' for ck in LINE_REG_CK)
            b.append(IND + '// register-staged line sweep: unit steps, natural layout, 32-bit indices')
            b.append(IND + 'DvmType CK = tune[DVMH_AMD_T_TILE_A];')
            b.append(IND + f'bool lreg = tune[DVMH_AMD_T_VARIANT] == 2 && (loopSteps[{dep}] == 1 || loopSteps[{dep}] == -1) && '
                     f'{X}_tfm == 0 && kernelIndexT == rt_INT && threads.x * threads.y * threads.z <= 256 && ({cks});')
            b.append(IND + 'if (lreg) {')
            b += [IND * 2 + s for s in line_reg_launch(L, dep)]
            b.append(IND + '} else {')
            b += [IND * 2 + s for s in line_launch(L, dep)]
            b.append(IND + '}')
        else:
            b += [IND + s for s in line_launch(L, dep)]
    # two or more dependent axes: hyperplane kernels for every mask with >= 2 bits
    for m in range(2, r + 1):
        for deps in itertools.combinations(range(r), m):
            mask = sum(1 << (r - 1 - a) for a in deps)
            b.append(f'}} else if (dependency_mask == {mask}) {{')
            if m == 3 and L.tile_eligible():
                X = L.across['array']
                xct = [p for p in L.arrays if p[1] == X][0][2]
                unit = ' && '.join(f'loopSteps[{i}] == 1' for i in range(3))
                b.append(IND + '// tiled variant: unit forward steps, natural layout, tile fits 64 KB of LDS')
                b.append(IND + f'bool tiled = tune[DVMH_AMD_T_VARIANT] == 1 && {unit} && {X}_tfm == 0 && threads.y == 1 && '
                         'threads.z == 1 && tune[DVMH_AMD_T_TILE_A] >= 2 && tune[DVMH_AMD_T_TILE_B] >= 2 && '
                         f'(tune[DVMH_AMD_T_TILE_A] + 2) * (tune[DVMH_AMD_T_TILE_A] + 2) * (tune[DVMH_AMD_T_TILE_B] + 2) * (DvmType)sizeof({xct}) <= 64512;')
                b.append(IND + 'if (tiled) {')
                b += [IND * 2 + s for s in tile_launch(L)]
                b.append(IND + '} else {')
                b += [IND * 2 + s for s in wave_launch(L, list(deps))]
                b.append(IND + '}')
            else:
                b += [IND + s for s in wave_launch(L, list(deps))]
    b.append('} else {')
    b.append(IND + f'fprintf(stderr, "{L.name}: unsupported dependency mask %ld\\n", (long)dependency_mask);')
    b.append(IND + 'abort();')
    b.append('}')
    h.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    h.append('}')
    return '\n'.join(h) + '\n'


def seq_kernel(L):
    """The sequential part on the device: one lane runs the statement."""
    fp, setup = kernel_array_params(L, unit=True)
    k = [f'template <typename IndexT>',
         f'__global__ __launch_bounds__(64) void {L.name}_kernel({", ".join(fp)}) {{',
         IND + 'if (blockIdx.x != 0 || threadIdx.x != 0) return;']
    k += [IND + x for x in setup]
    k.append(indent('{\n' + indent(L.body + '\n', 1) + '}', 1))
    k.append('}')
    return '\n'.join(k) + '\n'


def seq_handler(L):
    h = [f'extern "C" void {L.name}_cuda(DvmType *pLoopRef{"".join(", DvmType " + p[1] + "_hdr[]" if p[0] == "array" else ", " + p[2] + " *" + p[1] + "_ptr" for p in L.params)}) {{']
    b = ['DvmType tmpVar;',
         'DvmType loop_ref = *pLoopRef;',
         'DvmType device_num = dvmh_loop_get_device_num_C(loop_ref);']
    for p in L.arrays:
        _, nm, ct, ar = p
        b.append(f'{ct} *{nm} = ({ct} *)dvmh_get_natural_base_C(device_num, {nm}_hdr);')
        b.append(f'DvmType {nm}_devHdr[64 + 7];')
        b.append(f'tmpVar = dvmh_fill_header_C(device_num, {nm}, {nm}_hdr, {nm}_devHdr, 0);')
        b.append(f'if (tmpVar != 0) {{ fprintf(stderr, "{L.name}: array {nm} is not in natural layout\\n"); abort(); }}')
    for p in L.scalars:
        b.append(f'{p[2]} *{p[1]} = ({p[2]} *)dvmh_get_device_addr_C(device_num, {p[1]}_ptr);')
    b.append('hipStream_t stream;')
    b.append('dvmh_loop_cuda_get_config_C(loop_ref, 0, 0, 0, &stream, 0);')
    b.append('int kernelIndexT = (int)dvmh_loop_guess_index_type_C(loop_ref);')
    b.append('dvmh_amd_mark_kernel_C(loop_ref, 0);')
    for it, cond in (('int', 'kernelIndexT == rt_INT'), ('long long', 'kernelIndexT != rt_INT')):
        args = array_fact_args(L, it)
        b.append(f'if ({cond}) hipLaunchKernelGGL(({L.name}_kernel<{it}>), dim3(1), dim3(64), 0, stream, {", ".join(args)});')
    b.append('dvmh_amd_mark_kernel_C(loop_ref, 1);')
    h.append(indent('\n'.join(b) + '\n', 1).rstrip('\n'))
    h.append('}')
    return '\n'.join(h) + '\n'


def gpu_loop(L):
    text = []
    if L.sequential:
        return seq_kernel(L) + '\n' + seq_handler(L)
    if L.across:
        r = L.rank
        for dep in range(r):
            text.append(line_kernel(L, dep))
        for m in range(2, r + 1):
            for deps in itertools.combinations(range(r), m):
                text.append(wave_kernel(L, list(deps)))
        if L.tile_eligible():
            text.append(tile_kernel(L))
        if L.line_tile_eligible():
            text.append(line_tile_kernel(L))
            text.append(line_vec_kernel(L))
        for dep in range(r - 1):
            if L.line_stage_eligible(dep):
                text.append(line_reg_kernel(L, dep))
        text.append(across_handler(L))
        if L.chain_eligible():
            text.append(chain_handler(L))
    else:
        text.append(parallel_kernel(L))
        if L.stencil:
            text.append(march_kernel(L))
        text.append(parallel_handler(L))
        if L.batch:
            text.append(batch_kernel(L))
            if L.stencil and L.rank in (2, 3):
                text.append(batch_march_kernel(L))
            text.append(batch_handler(L))
    return '\n'.join(text)


# ----------------------------------------------------------------- driver
def emit(spec_path, outdir):
    with open(spec_path) as f:
        spec = ast.literal_eval(f.read())
    prog = spec['program']
    loops = [Loop(d, prog) for d in spec['loops']]
    inc = spec.get('include')
    incs = [inc] if isinstance(inc, str) else list(inc or [])
    defs = list(spec.get('defines', []))   # the program's own macros, which loop bodies may use
    host = ['/* Host handlers generated from ' + os.path.basename(spec_path) + ' by emit/hipgen.py. */',
            '#include <math.h>', '#include <stdio.h>', '#include <stdlib.h>',
            '#ifdef _OPENMP', '#include <omp.h>', '#endif', '#include "dvmhlib2.h"']
    host += [f'#include "{x}"' for x in incs] + defs
    host.append('')
    for L in loops:
        host.append(host_handler(L))
    hip = ['// GPU handlers generated from ' + os.path.basename(spec_path) + ' by emit/hipgen.py (gfx950).',
           '#include <hip/hip_runtime.h>', '#include <math.h>', '#include <stdio.h>', '#include <stdlib.h>', '#include <string.h>',
           '#include "dvmhlib2.h"', '#include "dvmh_amd.h"', '#include "dvmh_hip_device.h"']
    hip += [f'#include "{x}"' for x in incs] + defs
    hip.append('')
    for L in loops:
        if not L.host_only:
            hip.append(gpu_loop(L))
    # consecutive loops that may run fused: the runtime learns of each pair's
    # handler when the library is loaded and uses it when the two loops run
    # one after the other in a deferred region or a batch
    pairs = [(L1, L2) for L1, L2 in zip(loops, loops[1:]) if not L1.host_only and not L2.host_only and pair_eligible(L1, L2)]
    for L1, L2 in pairs:
        hip.append(pair_code(L1, L2))
    if pairs:
        hip.append('__attribute__((constructor)) static void dvmh_register_pairs(void) {')
        for L1, L2 in pairs:
            hip.append(f'    dvmh_amd_register_pair_C((DvmHandlerFunc){L1.name}_cuda, (DvmHandlerFunc){L2.name}_cuda, '
                       f'(DvmHandlerFunc){pair_name(L1, L2)}_cuda_pair);')
        hip.append('}')
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, f'{prog}_host.c'), 'w') as f:
        f.write('\n'.join(host))
    with open(os.path.join(outdir, f'{prog}_hip.hip'), 'w') as f:
        f.write('\n'.join(hip))


if __name__ == '__main__':
    if len(sys.argv) != 3:
        sys.exit('usage: hipgen.py <program.loops> <outdir>')
    emit(sys.argv[1], sys.argv[2])
