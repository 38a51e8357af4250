#!/usr/bin/env python3
"""C-DVMH front end: converts a .cdv source (C with `#pragma dvm` directives)
into

  <prog>.c      the converted host program: the same functions, with the
                directives replaced by calls of the DVMH C interface
                (`dvmhlib2.h`) in the shape cdvmh writes them
                (cdvmh/src/converter_main.cpp: distributed-array declarations
                :680-750, malloc/free of distributed arrays :3370-3400 and
                :3610-3645, host references to distributed elements
                :3407-3500, regions and their registrations :2990-3095,
                parallel loops and handler registration :2132-2401);
  <prog>.loops  the description of every parallel loop (loop variables and
                steps, parameters in registration order, reduction, ACROSS,
                private and shadow_renew clauses, the body text) that
                emit/hipgen.py turns into host (C/OpenMP) and HIP handlers.

Supported directives (the subset of the C-DVMH language the configuration
programs use, plus the obvious neighbours):
  array distribute[block]...[block] / align([i][j] with A[i][j])
        on a declaration of a pointer to array (`double (*A)[n][n];`, storage
        from a later `A = malloc(bytes);`) or of a complete array
        (`double A[N][M];`); shadow widths are 1;
  region [in(...)] [out(...)] [inout(...)] [local(...)] [async]
        followed by a compound statement; arrays not named in a clause are
        registered IN if the region's loops only read them, INOUT otherwise;
        scalars read by loop bodies are registered IN;
  parallel([i][j]... on A[i+c][j+c]...) with shadow_renew(A, ...),
        reduction(sum|product|max|min|and|or|xor(v), maxloc|minloc(v, loc)),
        across(A[lo:hi]...), private(v, ...), cuda_block(x[, y[, z]]),
        followed by a perfect nest of `for` loops of unit or constant step;
  actual(...) / get_actual(...) of scalars, C arrays, distributed arrays and
        distributed sub-arrays A[lo:hi]...;
  free(A) of a distributed array; element references A[i][j] in host code.

Anything else in the source is copied unchanged.  Constructs outside the
subset stop the conversion with a message naming the line.
"""
import os
import re
import sys

# ---------------------------------------------------------------- tokens
_TOKEN_RE = re.compile(r'''
    (?P<nl>\n)
  | (?P<ws>[ \t\r\f\v]+)
  | (?P<cmt>/\*.*?\*/|//[^\n]*)
  | (?P<pp>\#(?:[^\n\\]|\\.)*)
  | (?P<str>"(?:\\.|[^"\\\n])*")
  | (?P<chr>'(?:\\.|[^'\\\n])*')
  | (?P<num>0[xX][0-9a-fA-F]+[uUlL]*|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?[uUlLfF]*)
  | (?P<id>[A-Za-z_]\w*)
  | (?P<op>->|\+\+|--|<<=|>>=|<<|>>|<=|>=|==|!=|&&|\|\||[-+*/%&|^]=|[{}()\[\];,.?:~!<>=+\-*/%&|^])
''', re.S | re.X)

TYPE_WORDS = {'void', 'char', 'short', 'int', 'long', 'float', 'double', 'signed', 'unsigned', 'const', 'static',
              'volatile', 'register', 'extern', '_Bool', 'size_t'}
QUALIFIERS = {'const', 'static', 'volatile', 'register', 'extern'}

RT_TYPE = {'char': 'rt_CHAR', 'unsigned char': 'rt_UCHAR', 'int': 'rt_INT', 'unsigned int': 'rt_UINT',
           'unsigned': 'rt_UINT', 'long': 'rt_LONG', 'unsigned long': 'rt_ULONG', 'long long': 'rt_LLONG',
           'unsigned long long': 'rt_ULLONG', 'float': 'rt_FLOAT', 'double': 'rt_DOUBLE'}
RED_FUNC = {'sum': 'rf_SUM', 'product': 'rf_PROD', 'prod': 'rf_PROD', 'max': 'rf_MAX', 'min': 'rf_MIN',
            'and': 'rf_AND', 'or': 'rf_OR', 'xor': 'rf_XOR', 'maxloc': 'rf_MAXLOC', 'minloc': 'rf_MINLOC'}


class ConvertError(Exception):
    pass


class Tok:
    __slots__ = ('kind', 'text', 'start', 'end', 'line')

    def __init__(self, kind, text, start, end, line):
        self.kind, self.text, self.start, self.end, self.line = kind, text, start, end, line

    def __repr__(self):
        return f'{self.kind}:{self.text!r}@{self.line}'


def tokenize(src):
    toks, pos, line = [], 0, 1
    while pos < len(src):
        m = _TOKEN_RE.match(src, pos)
        if not m:
            raise ConvertError(f'line {line}: cannot tokenize {src[pos:pos + 20]!r}')
        kind = m.lastgroup
        text = m.group()
        if kind not in ('nl', 'ws', 'cmt'):
            toks.append(Tok(kind, text, pos, m.end(), line))
        line += text.count('\n')
        pos = m.end()
    return toks


def matching(toks, i):
    """Index of the bracket closing toks[i]."""
    pairs = {'(': ')', '[': ']', '{': '}'}
    open_, close = toks[i].text, pairs[toks[i].text]
    depth = 0
    for j in range(i, len(toks)):
        if toks[j].text == open_ and toks[j].kind == 'op':
            depth += 1
        elif toks[j].text == close and toks[j].kind == 'op':
            depth -= 1
            if depth == 0:
                return j
    raise ConvertError(f'line {toks[i].line}: unbalanced {open_!r}')


def split_top(text, sep=','):
    """Split at `sep` outside brackets."""
    out, depth, cur = [], 0, ''
    for ch in text:
        if ch in '([{':
            depth += 1
        elif ch in ')]}':
            depth -= 1
        if ch == sep and depth == 0:
            out.append(cur.strip())
            cur = ''
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def brackets(text):
    """'[a][b:c]' -> ['a', 'b:c'] (top-level square brackets)."""
    out, depth, cur = [], 0, ''
    for ch in text.strip():
        if ch == '[':
            if depth > 0:
                cur += ch
            depth += 1
        elif ch == ']':
            depth -= 1
            if depth > 0:
                cur += ch
            else:
                out.append(cur.strip())
                cur = ''
        elif depth > 0:
            cur += ch
    return out


# ---------------------------------------------------------------- directives
def parse_clauses(text):
    """'on A[i][j]) shadow_renew(A), reduction(max(e))' style clause lists:
    returns [(name, argument text)] for name(args) items."""
    out, i = [], 0
    while i < len(text):
        m = re.compile(r'\s*,?\s*([A-Za-z_]\w*)\s*').match(text, i)
        if not m:
            break
        name = m.group(1)
        i = m.end()
        if i < len(text) and text[i] == '(':
            depth, j = 0, i
            while j < len(text):
                if text[j] == '(':
                    depth += 1
                elif text[j] == ')':
                    depth -= 1
                    if depth == 0:
                        break
                j += 1
            out.append((name, text[i + 1:j]))
            i = j + 1
        else:
            out.append((name, ''))
    return out


class Directive:
    def __init__(self, tok):
        self.line = tok.line
        m = re.match(r'#\s*pragma\s+dvm\s+(\w+)\s*(.*)$', tok.text.replace('\\\n', ' '), re.S)
        self.kind = m.group(1)
        self.rest = m.group(2).strip()


# ---------------------------------------------------------------- symbols
class Var:
    def __init__(self, name, ctype, dims=None, pointer_rank=0, is_param=False):
        self.name = name
        self.ctype = ctype              # base element type text, e.g. 'double', 'unsigned long long'
        self.dims = dims or []          # declared extents of a C array / pointer-to-array tail
        self.pointer_rank = pointer_rank
        self.is_param = is_param
        self.dvm = None                 # DvmArray when distributed


class DvmArray:
    def __init__(self, var, rank, sizes, incomplete, rule):
        self.var = var
        self.rank = rank
        self.sizes = sizes              # extents as text (first may be None for incomplete)
        self.incomplete = incomplete
        self.rule = rule                # ('distribute', [...]) or ('align', dummy vars, target, target subscripts)
        self.shadow = [('1', '1')] * rank   # shadow widths (low, high) per axis
        self.line = 0


def parse_declaration(text):
    """'double (*A)[n][n]' / 'double A[10][20]' / 'long n = 5, m' / 'int i, j' ->
    [(name, base type, dims, pointer-to-array rank, initialiser)]"""
    text = text.strip()
    words = []
    m = re.match(r'((?:(?:' + '|'.join(sorted(TYPE_WORDS, key=len, reverse=True)) + r')\b\s*)+)', text)
    if not m:
        return None
    base = ' '.join(w for w in m.group(1).split() if w not in QUALIFIERS)
    rest = text[m.end():]
    out = []
    for d in split_top(rest):
        init = None
        if '=' in d:
            eq = d.index('=')
            if d[eq + 1:eq + 2] != '=':
                d, init = d[:eq].strip(), d[eq + 1:].strip()
        pm = re.match(r'\(\s*\*\s*([A-Za-z_]\w*)\s*\)\s*((?:\[[^\]]*\]\s*)*)$', d)
        if pm:
            out.append((pm.group(1), base, brackets(pm.group(2)), 1, init))
            continue
        am = re.match(r'(\**)\s*([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)*)$', d)
        if not am:
            return None
        ptr = len(am.group(1))
        name = am.group(2)
        out.append((name, base + ('*' * ptr if ptr else ''), brackets(am.group(3)), 0, init))
    return out


# ---------------------------------------------------------------- conversion
class Converter:
    def __init__(self, src, filename, prog):
        self.src = src
        self.file = filename
        self.prog = prog
        self.toks = tokenize(src)
        self.loops = []            # loop descriptions for hipgen
        self.handlers = []         # (name, param kinds) for prototypes
        self.region_no = 0
        self.includes = []
        self.defines = []          # file-scope #define lines, repeated in the handler files
        self.globals = {}          # file-scope variables and distributed arrays
        self.global_arrays = []
        self.inherits = {}         # function -> positions of its inherited (distributed array) parameters

    # -- helpers
    def text(self, a, b):
        """Source text of tokens [a, b)."""
        if a >= b:
            return ''
        return self.src[self.toks[a].start:self.toks[b - 1].end]

    def err(self, tok_or_line, msg):
        line = tok_or_line if isinstance(tok_or_line, int) else tok_or_line.line
        raise ConvertError(f'{self.file}:{line}: {msg}')

    def line_call(self, line):
        return f'dvmh_line_C({line}, "{self.file}");'

    # -- top level
    def convert(self):
        out, i, last = [], 0, 0
        toks = self.toks
        stmt = 0            # first token of the current file-scope statement
        pending = None      # array directive waiting for its file-scope declaration
        inherit = None      # inherit directive waiting for its function declaration
        while i < len(toks):
            t = toks[i]
            if t.kind == 'pp':
                if re.match(r'#\s*include', t.text):
                    self.includes.append(t.text.strip())
                if re.match(r'#\s*define\b', t.text):
                    self.defines.append(t.text.strip())   # loop bodies may use the program's macros
                if re.match(r'#\s*pragma\s+dvm\b', t.text):
                    d = Directive(t)
                    if d.kind not in ('array', 'inherit'):
                        self.err(t, f'directive {d.kind!r} outside functions is not supported')
                    if pending or inherit:
                        self.err(t, 'one directive per declaration')
                    if d.kind == 'inherit':
                        m = re.match(r'\((.*)\)$', d.rest, re.S)
                        if not m:
                            self.err(t, 'inherit(names) expected')
                        inherit = (t.line, [x.strip() for x in split_top(m.group(1))])
                    else:
                        pending = d
                    out.append(self.src[last:t.start])
                    last = t.end
                i += 1
                stmt = i
                continue
            if t.text == ';' and t.kind == 'op':
                # a file-scope declaration: globals are visible to every function
                text = self.text(stmt, i)
                if inherit:
                    # a prototype of a function with inherited parameters
                    out.append(self.src[last:toks[stmt].start])
                    out.append(self.inherit_header(stmt, i, inherit)[1])
                    last = toks[i - 1].end
                    inherit = None
                elif pending:
                    v, a = self.array_spec(pending, text)
                    self.globals[v.name] = v
                    self.global_arrays.append(a)
                    out.append(self.src[last:toks[stmt].start])
                    out.append(f'DvmType {v.name}[64];')
                    last = t.end
                    pending = None
                else:
                    for nm, ct, dims, pr, _ in (parse_declaration(re.sub(r'^\s*(static|extern)\s+', '', text)) or []):
                        self.globals[nm] = Var(nm, ct, dims, pr)
                i += 1
                stmt = i
                continue
            if t.text == '{' and t.kind == 'op' and i > 0 and toks[i - 1].text == '=':
                i = matching(toks, i) + 1    # the initialiser of a file-scope variable
                continue
            if t.text == '{' and t.kind == 'op':
                if pending:
                    self.err(pending.line, 'an array directive at file scope is followed by a declaration')
                # a function definition: `... name ( params ) {`
                j = i - 1
                if j < 0 or toks[j].text != ')':
                    self.err(t, 'only function definitions may open a block at file scope')
                k = j
                depth = 0
                while k >= 0:
                    if toks[k].text == ')':
                        depth += 1
                    elif toks[k].text == '(':
                        depth -= 1
                        if depth == 0:
                            break
                    k -= 1
                name_tok = toks[k - 1]
                close = matching(toks, i)
                inherited = ()
                if inherit:
                    out.append(self.src[last:toks[stmt].start])
                    inherited, head = self.inherit_header(stmt, i, inherit)
                    out.append(head + self.src[toks[i - 1].end:t.start])
                    inherit = None
                else:
                    out.append(self.src[last:t.start])
                out.append(self.function(name_tok.text, k, j, i, close, inherited))
                last = toks[close].end
                i = close + 1
                stmt = i
                continue
            i += 1
        if pending:
            self.err(pending.line, 'array directive without a declaration')
        if inherit:
            self.err(inherit[0], 'inherit directive without a function declaration')
        out.append(self.src[last:])
        if self.global_arrays:
            # the file-scope arrays' DVM objects, created once by the first
            # converted function that runs, in declaration order
            self.scope = dict(self.globals)
            body = ''.join('    ' + self.array_create(a, '    ') + '\n' for a in self.global_arrays)
            out.append(f'\nstatic void {self.prog}_globals_init(void)\n{{\n    static int done = 0;\n'
                       f'    if (done) return;\n    done = 1;\n{body}}}\n')
        return ''.join(out)

    def inherit_header(self, a, b, inherit):
        """Declaration tokens [a, b) of a function under `inherit(names)`:
        (the inherited parameter names, the declaration with each of them
        turned into the DvmType header array its caller passes).  The
        reference converter rewrites such a parameter the same way
        (dvm_sys/cdvmh/src/converter_main.cpp:856-863); every name must be a
        parameter (:920-928)."""
        toks = self.toks
        line, names = inherit
        r = b - 1
        if toks[r].text != ')':
            self.err(line, 'inherit directive not followed by a function declaration')
        k = r
        depth = 0
        while k >= a:
            if toks[k].text == ')':
                depth += 1
            elif toks[k].text == '(':
                depth -= 1
                if depth == 0:
                    break
            k -= 1
        fname = toks[k - 1].text
        if fname == 'main':
            self.err(line, 'main cannot inherit distributed arrays')
        text, base = self.text(a, b), toks[a].start
        pos, found = [], []
        for n, (p, q) in reversed(list(enumerate(self.param_ranges(k + 1, r)))):
            d = parse_declaration(self.text(p, q)) if q > p else None
            if d and d[0][0] in names:
                text = text[:toks[p].start - base] + f'DvmType {d[0][0]}[64]' + text[toks[q - 1].end - base:]
                pos.insert(0, n)
                found.insert(0, d[0][0])
        missing = [x for x in names if x not in found]
        if missing:
            self.err(line, f'inherited names that are not parameters of {fname}: {", ".join(missing)}')
        prev = self.inherits.setdefault(fname, pos)
        if prev != pos:
            self.err(line, f'{fname}: inherited parameters differ from its earlier declaration')
        return found, text

    def param_ranges(self, a, b):
        """Token ranges [p, q) of the comma-separated items within [a, b)."""
        out, p, depth = [], a, 0
        for k in range(a, b):
            x = self.toks[k].text
            if x in ('(', '[', '{'):
                depth += 1
            elif x in (')', ']', '}'):
                depth -= 1
            elif x == ',' and depth == 0:
                out.append((p, k))
                p = k + 1
        if b > a:
            out.append((p, b))
        return out

    def empty_macros(self):
        """Function-like macros the program defines as nothing (`#define DVM(dir)`,
        the old C-DVM directive form, which cdvmh leaves to the preprocessor):
        an invocation of one is not C; it is skipped where statements are split
        and dropped where declarations are read."""
        out = []
        for dfn in self.defines:
            m = re.match(r'#\s*define\s+([A-Za-z_]\w*)\(([^)]*)\)\s*$', dfn)
            if m:
                out.append(m.group(1))
        return out

    def header_macros(self):
        """Function-like macros whose replacement is a loop or if header
        (`#define DO(v,l,h,s) for(v=(l); v<=(h); v+=(s))`): an invocation heads
        the statement that follows it."""
        out = []
        for dfn in self.defines:
            m = re.match(r'#\s*define\s+([A-Za-z_]\w*)\([^)]*\)\s*(for|while|if)\s*\(', dfn)
            if m:
                out.append(m.group(1))
        return out

    def macro_head(self, i, b):
        """Index of the ')' closing the invocation at toks[i] of an empty or
        header macro, else None."""
        t = self.toks[i]
        if t.kind == 'id' and i + 1 < b and self.toks[i + 1].text == '(' and \
                (t.text in self.empty_macros() or t.text in self.header_macros()):
            return matching(self.toks, i + 1)
        return None

    def strip_empty_macros(self, text):
        for name in self.empty_macros():
            while True:
                m = re.search(r'\b' + name + r'\s*\(', text)
                if not m:
                    break
                depth, k = 0, m.end() - 1
                while k < len(text):
                    depth += {'(': 1, ')': -1}.get(text[k], 0)
                    if depth == 0:
                        break
                    k += 1
                text = text[:m.start()] + text[k + 1:]
        return text

    def function(self, name, lparen, rparen, lbrace, rbrace, inherited=()):
        self.scope = dict(self.globals)
        for p in split_top(self.text(lparen + 1, rparen)):
            p = self.strip_empty_macros(p).strip()
            if p in ('', 'void'):
                continue
            d = parse_declaration(p)
            if not d:
                self.err(self.toks[lparen], f'cannot read parameter {p!r}')
            nm, ct, dims, pr, _ = d[0]
            v = Var(nm, ct, dims, pr, is_param=True)
            if dims and not pr:           # `double x[]` parameter: a pointer
                v.pointer_rank = 0
            if nm in inherited:
                # the caller's distributed array, passed as its header: the
                # extents come from the header, the element type and rank
                # from the declaration (`double (*a)[n][m]`, `double a[][m]`)
                rank = len(dims) + (1 if pr else 0)
                if rank < 1:
                    self.err(self.toks[lparen], f'inherited parameter {nm} is not an array')
                v.dvm = DvmArray(v, rank, ([None] if pr else []) + list(dims), True, None)
                v.dvm.line = self.toks[lparen].line
            self.scope[nm] = v
        self.arrays_here = []
        body = self.block(lbrace + 1, rbrace, ctx={'region': None})
        ind = '    '
        head = '{\n' + ind + 'dvmh_init_lib_C(0);'
        if self.global_arrays:
            head += f'\n{ind}{self.prog}_globals_init();'
        tail = ''
        for a in self.arrays_here:
            tail += ind + ('dvmh_forget_header_' if a.incomplete else 'dvmh_delete_object_') + f'({a.var.name});\n'
        return head + body + tail + '}'

    # -- statements
    def block(self, a, b, ctx):
        """Convert the statements of tokens [a, b) (inside a block)."""
        out, i = [], a
        pending = []      # directives waiting for their statement
        prev_end = self.toks[a - 1].end if a > 0 else 0
        while i < b:
            t = self.toks[i]
            gap = self.src[prev_end:t.start]
            if t.kind == 'pp' and re.match(r'#\s*pragma\s+dvm\b', t.text):
                d = Directive(t)
                out.append(gap)
                if d.kind in ('actual', 'get_actual'):
                    out.append(self.actual(d, self.indent_of(i)))
                elif d.kind == 'realign':
                    out.append(self.realign(d, self.indent_of(i)))
                elif d.kind == 'redistribute':
                    out.append(self.redistribute(d, self.indent_of(i)))
                elif d.kind in ('array', 'region', 'parallel', 'remote_access'):
                    pending.append(d)
                else:
                    self.err(t, f'directive {d.kind!r} is not supported')
                prev_end = t.end
                i += 1
                continue
            j = self.statement_end(i, b)
            out.append(gap)
            if pending:
                d = pending.pop(0)
                if pending:
                    self.err(d.line, 'one directive per statement')
                if d.kind == 'array':
                    out.append(self.array_decl(d, i, j))
                elif d.kind == 'region':
                    out.append(self.region(d, i, j, ctx))
                elif d.kind == 'remote_access':
                    out.append(self.remote_block(d, i, j, ctx))
                else:
                    out.append(self.parallel(d, i, j, ctx))
            else:
                out.append(self.plain_statement(i, j, ctx))
            prev_end = self.toks[j - 1].end
            i = j
        if pending:
            self.err(pending[0].line, 'directive without a statement')
        out.append(self.src[prev_end:self.toks[b].start] if b < len(self.toks) else '')
        return ''.join(out)

    def statement_end(self, i, b):
        """Index one past the statement starting at toks[i]."""
        t = self.toks[i]
        if t.text == '{':
            return matching(self.toks, i) + 1
        if t.kind == 'pp':
            return i + 1
        k = self.macro_head(i, b)
        if k is not None:
            return self.statement_end(k + 1, b)
        if t.text in ('for', 'while', 'switch'):
            k = matching(self.toks, i + 1)
            return self.statement_end(k + 1, b)
        if t.text == 'if':
            k = matching(self.toks, i + 1)
            e = self.statement_end(k + 1, b)
            if e < b and self.toks[e].text == 'else':
                e = self.statement_end(e + 1, b)
            return e
        if t.text == 'do':
            e = self.statement_end(i + 1, b)      # body
            k = matching(self.toks, e + 1)        # while ( ... )
            return k + 2
        depth = 0
        j = i
        while j < b:
            x = self.toks[j].text
            if x in '([{' and self.toks[j].kind == 'op':
                depth += 1
            elif x in ')]}' and self.toks[j].kind == 'op':
                depth -= 1
            elif x == ';' and depth == 0:
                return j + 1
            j += 1
        self.err(t, 'statement without ";"')

    def plain_statement(self, i, j, ctx):
        toks = self.toks
        t = toks[i]
        if t.text == '{':
            k = matching(toks, i)
            return '{' + self.block(i + 1, k, ctx) + '}'
        k = self.macro_head(i, j)
        if k is not None and k + 1 < j:
            # DVM(...) or a loop-header macro: kept as written, the statement after it converted
            return self.host_expr(i, k + 1, ctx) + self.src[toks[k].end:toks[k + 1].start] + self.plain_statement(k + 1, j, ctx)
        if t.text in ('for', 'while', 'switch', 'if'):
            k = matching(toks, i + 1)
            head = self.host_expr(i, k + 1, ctx)
            e = self.statement_end(k + 1, j)
            body = self.plain_statement(k + 1, e, ctx)
            res = head + self.src[toks[k].end:toks[k + 1].start] + body
            if t.text == 'if' and e < j and toks[e].text == 'else':
                res += self.src[toks[e - 1].end:toks[e].start] + 'else' + \
                    self.src[toks[e].end:toks[e + 1].start] + self.plain_statement(e + 1, j, ctx)
            return res
        if t.text == 'do':
            e = self.statement_end(i + 1, j)
            return 'do' + self.src[t.end:toks[i + 1].start] + self.plain_statement(i + 1, e, ctx) + \
                self.src[toks[e - 1].end:toks[e].start] + self.host_expr(e, j, ctx)
        if t.text == 'return' and self.arrays_here:
            # the function's distributed arrays are released first (as at its
            # end); the returned value must not read them
            if self.uses_dvm_array(i + 1, j):
                self.err(t, 'a return value that reads a distributed array is not supported')
            ind = self.indent_of(i)
            rel = ''.join(('dvmh_forget_header_' if a.incomplete else 'dvmh_delete_object_') + f'({a.var.name}); '
                          for a in self.arrays_here)
            return '{ ' + rel + self.host_expr(i, j, ctx) + ' }'
        # declarations: remember the types
        d = parse_declaration(self.text(i, j - 1))
        if d:
            for nm, ct, dims, pr, _ in d:
                self.scope[nm] = Var(nm, ct, dims, pr)
            return self.host_expr(i, j, ctx)
        # A = malloc(...) / free(A) of distributed arrays
        m = re.match(r'\s*([A-Za-z_]\w*)\s*=\s*(?:\([^;]*?\)\s*)?(malloc|calloc)\s*\((.*)\)\s*;\s*$', self.text(i, j), re.S)
        if m and m.group(1) in self.scope and self.scope[m.group(1)].dvm:
            a = self.scope[m.group(1)].dvm
            args = split_top(m.group(3))
            nbytes = args[0] if m.group(2) == 'malloc' else f'({args[0]}) * ({args[1]})'
            return self.line_call(t.line) + '\n' + self.indent_of(i) + f'dvmh_array_alloc_C({a.var.name}, {nbytes});' + \
                '\n' + self.indent_of(i) + self.distribution_call(a)
        m = re.match(r'\s*free\s*\(\s*([A-Za-z_]\w*)\s*\)\s*;\s*$', self.text(i, j))
        if m and m.group(1) in self.scope and self.scope[m.group(1)].dvm:
            return self.line_call(t.line) + '\n' + self.indent_of(i) + f'dvmh_array_free_C({m.group(1)});'
        if ctx['region'] is not None and self.uses_dvm_array(i, j):
            return self.sequential_part(i, j, ctx)
        return self.host_expr(i, j, ctx)

    def sequential_part(self, i, j, ctx):
        """A region's statement outside parallel loops that uses distributed
        arrays (cdvmh/src/converter_main.cpp:1310-1330): a loop of rank 0 whose
        handlers run the statement once where the region runs."""
        t = self.toks[i]
        line0 = self.src.rfind('\n', 0, t.start) + 1
        body = self.dedent(self.src[line0:self.toks[j - 1].end])
        params = []
        for x in tokenize(body):
            if x.kind != 'id' or x.text not in self.scope or any(p[1] == x.text for p in params):
                continue
            v = self.scope[x.text]
            if v.dvm:
                params.append(('array', x.text, v.ctype, v.dvm.rank))
            elif not v.dims and not v.pointer_rank:
                params.append(('scalar', x.text, v.ctype))
            else:
                self.err(t, f'{x.text}: C arrays and pointers in a region\'s sequential part are not supported')
        name = f'sequence_{self.prog}_{t.line}'
        while any(L['name'] == name for L in self.loops):
            name += '_'
        self.loops.append({'name': name, 'vars': [], 'steps': [], 'params': params, 'body': '\n' + body + '\n',
                           'sequential': True})
        self.handlers.append((name, params, False))
        for p in params:
            written = re.search(r'\b' + p[1] + r'\s*(?:\[[^;]*?\]\s*)*(?:=(?!=)|\+=|-=|\*=|/=|\+\+|--)', body) or \
                re.search(r'(?:\+\+|--)\s*' + p[1] + r'\b', body)
            if p[0] == 'array':
                ctx['arrays'][p[1]] = ctx['arrays'].get(p[1], False) or bool(written)
            else:
                ctx['scalars'][p[1]] = p[2]
                if written:
                    ctx.setdefault('scalars_out', set()).add(p[1])
        lp = ctx['loop']
        pargs = ', '.join(p[1] if p[0] == 'array' else '&' + p[1] for p in params)
        out = [self.line_call(t.line),
               f'{lp} = dvmh_loop_create_C({ctx["region"]}, 0);',
               f'dvmh_loop_register_handler_C({lp}, DEVICE_TYPE_HOST, 0, '
               f'dvmh_handler_func_C((DvmHandlerFunc){name}_host, {len(params)}, {pargs}));',
               f'dvmh_loop_register_handler_C({lp}, DEVICE_TYPE_CUDA, 0, '
               f'dvmh_handler_func_C((DvmHandlerFunc){name}_cuda, {len(params)}, {pargs}));',
               f'dvmh_loop_perform_C({lp});']
        return ('\n' + self.indent_of(i)).join(out)

    def indent_of(self, i):
        s = self.src
        p = self.toks[i].start
        q = s.rfind('\n', 0, p) + 1
        return re.match(r'[ \t]*', s[q:p]).group()

    def uses_dvm_array(self, i, j):
        return any(t.kind == 'id' and t.text in self.scope and self.scope[t.text].dvm for t in self.toks[i:j])

    def host_expr(self, i, j, ctx):
        """Source of tokens [i, j) with host references A[..]..[..] to
        distributed arrays replaced by element accesses."""
        toks = self.toks
        out, k, last = [], i, toks[i].start
        headers = set()     # arguments passed to inherited parameters: the header itself
        while k < j:
            t = toks[k]
            if t.kind == 'id' and t.text in self.inherits and k + 1 < j and toks[k + 1].text == '(':
                args = self.param_ranges(k + 2, matching(toks, k + 1))
                for n in self.inherits[t.text]:
                    if n < len(args):
                        p, q = args[n]
                        w = self.scope.get(toks[p].text) if q == p + 1 and toks[p].kind == 'id' else None
                        if not (w and w.dvm):
                            self.err(t, f'{t.text}: argument {n + 1} must be a distributed array')
                        headers.add(p)
            v = self.scope.get(t.text) if t.kind == 'id' else None
            if k in headers:
                k += 1
                continue
            if v and v.dvm and k + 1 < j and toks[k + 1].text == '[' and (k == i or toks[k - 1].text not in ('.', '->')):
                subs, e = [], k + 1
                while e < j and toks[e].text == '[' and len(subs) < v.dvm.rank:
                    c = matching(toks, e)
                    subs.append(self.host_expr(e + 1, c, ctx) if c > e + 1 else '')
                    e = c + 1
                if len(subs) != v.dvm.rank:
                    self.err(t, f'{t.text}: host reference with {len(subs)} of {v.dvm.rank} subscripts')
                out.append(self.src[last:t.start])
                out.append(f'(*({v.ctype} *)dvmh_get_element_addr_C({t.text}, {v.dvm.rank}' +
                           ''.join(f', (DvmType)({s})' for s in subs) + '))')
                last = toks[e - 1].end
                k = e
                continue
            if v and v.dvm and (k == i or toks[k - 1].text not in ('.', '->')):
                # the whole array as a value (fwrite(B, ...), a call argument):
                # its host copy in natural layout (made actual by get_actual)
                out.append(self.src[last:t.start])
                out.append(f'((void *)dvmh_get_natural_base_C(0, {t.text}))')
                last = t.end
            k += 1
        out.append(self.src[last:toks[j - 1].end])
        return ''.join(out)

    def const_extent(self, dims):
        """The element count of a one-dimensional C array whose extent is an
        integer literal or a macro defined as one; None otherwise."""
        if len(dims) != 1 or not dims[0]:
            return None
        text = dims[0].strip()
        for _ in range(8):
            if re.fullmatch(r'\d+', text):
                return int(text)
            m = None
            for dfn in self.defines:
                m = re.match(r'#\s*define\s+' + re.escape(text) + r'\s+\(?\s*(\w+)\s*\)?\s*$', dfn)
                if m:
                    break
            if not m:
                return None
            text = m.group(1)
        return None

    # -- directives
    def array_spec(self, d, text):
        """The distributed array declared by `text` under array directive d:
        distribute[block|]... or align(...), optionally shadow[lo:hi]... ."""
        decl = parse_declaration(text)
        if not decl or len(decl) != 1:
            self.err(d.line, 'a distributed array declaration declares one array')
        nm, ct, dims, pr, init = decl[0]
        if init:
            self.err(d.line, 'a distributed array declaration has no initialiser')
        rank = len(dims) + (1 if pr else 0)
        if ct not in RT_TYPE:
            self.err(d.line, f'element type {ct!r} is not supported')
        v = Var(nm, ct, dims, pr)
        sizes = ([None] + dims) if pr else dims
        rule, shadow = None, None
        for cl in split_top(d.rest):
            cl = cl.strip()
            m = re.match(r'distribute\s*((?:\[[^\]]*\]\s*)+)$', cl)
            if m:
                fm = self.dist_formats(brackets(m.group(1)), rank, d.line)
                rule = ('distribute', fm)
                continue
            m2 = re.match(r'align\s*\(\s*((?:\[[^\]]*\]\s*)+)\s*with\s+([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)+)\)\s*$', cl)
            if m2:
                rule = ('align', brackets(m2.group(1)), m2.group(2), brackets(m2.group(3)))
                continue
            m3 = re.match(r'shadow\s*((?:\[[^\]]*\]\s*)+)$', cl)
            if m3:
                ws = brackets(m3.group(1))
                if len(ws) != rank:
                    self.err(d.line, 'shadow gives a width per axis')
                shadow = []
                for w in ws:
                    lo, sep, hi = w.partition(':')
                    shadow.append((lo.strip(), (hi if sep else lo).strip()))
                continue
            if cl:
                self.err(d.line, f'array directive clause {cl!r} is not supported')
        # no rule: distributed later by redistribute, or aligned by realign
        a = DvmArray(v, rank, sizes, bool(pr), rule)
        if shadow:
            a.shadow = shadow
        a.line = d.line
        v.dvm = a
        return v, a

    def array_create(self, a, ind):
        """Calls that create (complete) or declare (pointer) the array's DVM object."""
        nm, rank = a.var.name, a.rank
        typ = f'-{RT_TYPE[a.var.ctype]}'
        out = self.line_call(a.line) + '\n' + ind
        def w(x):
            return f'(DvmType){x}' if x.isdigit() else f'(DvmType)({x})'
        if a.incomplete:
            axes = ''.join(f', (DvmType)({s if s is not None else 0}), {w(lo)}, {w(hi)}' for s, (lo, hi) in zip(a.sizes, a.shadow))
            out += f'dvmh_array_declare_C({nm}, {rank}, {typ}{axes});'
        else:
            axes = ''.join(f', (DvmType)({s}), {w(lo)}, {w(hi)}' for s, (lo, hi) in zip(a.sizes, a.shadow))
            out += f'dvmh_array_create_C({nm}, {rank}, {typ}{axes});\n' + ind + self.distribution_call(a)
        return out

    def array_decl(self, d, i, j):
        v, a = self.array_spec(d, self.text(i, j - 1))
        self.scope[v.name] = v
        self.arrays_here.append(a)
        ind = self.indent_of(i)
        return f'DvmType {v.name}[64];\n' + ind + self.array_create(a, ind)

    def dist_formats(self, items, rank, line):
        """Per-axis distribution formats of distribute[...] / redistribute:
        block, [*] or [] (not distributed), wgtblock(w, n), genblock(g),
        multblock(m)."""
        fm = []
        for f in items:
            f = f.strip()
            if f in ('*', ''):
                fm.append(('', None))
                continue
            if f == 'block':
                fm.append(('block', None))
                continue
            m = re.match(r'(wgtblock|genblock|multblock)\s*\((.*)\)$', f, re.S)
            if not m:
                self.err(line, f'distribution format {f!r} is not supported')
            fm.append((m.group(1), [x.strip() for x in split_top(m.group(2))]))
        if len(fm) != rank:
            self.err(line, f'{len(fm)} distribution formats for an array of rank {rank}')
        return fm

    def dist_args(self, fm, line):
        """The distribution rule calls of `fm`; processor axes count the
        distributed array axes only (cdvmh/src/converter_main.cpp:530-590)."""
        out, mps = [], 1
        for f, args in fm:
            if f == '':
                out.append('dvmh_distribution_replicated_()')
                continue
            if f == 'block':
                out.append(f'dvmh_distribution_block_C({mps})')
            elif f == 'multblock':
                out.append(f'dvmh_distribution_multblock_C({mps}, {args[0]})')
            else:
                w = self.scope.get(args[0])
                if not w or w.ctype not in RT_TYPE or w.dvm:
                    self.err(line, f'{f}: {args[0]!r} is not an array of a known element type')
                if f == 'wgtblock':
                    if len(args) != 2:
                        self.err(line, 'wgtblock(weights, count) expected')
                    out.append(f'dvmh_distribution_wgtblock_C({mps}, {RT_TYPE[w.ctype]}, {args[0]}, {args[1]})')
                else:
                    out.append(f'dvmh_distribution_genblock_C({mps}, {RT_TYPE[w.ctype]}, {args[0]})')
            mps += 1
        return ', '.join(out)

    def distribution_call(self, a):
        nm = a.var.name
        if a.rule is None:
            return ''
        if a.rule[0] == 'distribute':
            return f'dvmh_distribute_C({nm}, {a.rank}, {self.dist_args(a.rule[1], a.line)});'
        _, dummies, target, tsubs = a.rule
        tv = self.scope.get(target)
        if not tv or not tv.dvm:
            self.err(0, f'align target {target!r} is not a distributed array')
        rules = []
        for s in tsubs:
            rules.append(self.linear_rule(s, dummies, 0))
        return f'dvmh_align_C({nm}, {target}, {tv.dvm.rank}, {", ".join(rules)});'

    def linear_rule(self, sub, names, line):
        """An alignment or mapping subscript over the dummy / loop names ->
        its rule call (cdvmh/src/converter_main.cpp:598-622): a*v + b gives
        (axis of v, a, b); an expression of none of the names a constant
        (0, 0, expr); an empty subscript replication (-1, 0, 0)."""
        sub = sub.strip()
        if not sub:
            return 'dvmh_alignment_linear_C(-1, 0, 0)'
        used = [t.text for t in tokenize(sub) if t.kind == 'id' and t.text in names]
        if not used:
            return f'dvmh_alignment_linear_C(0, 0, {sub})'
        m = re.match(r'^(?:(\w+)\s*\*\s*)?([A-Za-z_]\w*)\s*(?:([+-])\s*(\w+))?$', sub)
        if m and m.group(2) in names and (m.group(1) is None or m.group(1) not in names):
            mult = m.group(1) or '1'
            off = '0' if not m.group(3) else (m.group(4) if m.group(3) == '+' else '-' + m.group(4))
            return f'dvmh_alignment_linear_C({names.index(m.group(2)) + 1}, {mult}, {off})'
        self.err(line, f'subscript {sub!r} is not of the form a * v + b, a constant or empty')

    def realign(self, d, ind):
        """realign(X[i]... with T[...]... [new]) (cdvmh/src/converter_main.cpp:98-118)."""
        m = re.match(r'\(\s*([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)+)\s*with\s+([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)+)(\s*new)?\s*\)\s*$',
                     d.rest, re.S)
        if not m:
            self.err(d.line, 'realign(A[i]... with B[...]...) expected')
        a, b = self.scope.get(m.group(1)), self.scope.get(m.group(3))
        if not a or not a.dvm or not b or not b.dvm:
            self.err(d.line, 'realign relates two distributed arrays')
        dummies = [x.strip() for x in brackets(m.group(2))]
        if len(dummies) != a.dvm.rank:
            self.err(d.line, f'{a.name}: {len(dummies)} of {a.dvm.rank} dummy subscripts')
        tsubs = brackets(m.group(4))
        if len(tsubs) != b.dvm.rank:
            self.err(d.line, f'{b.name}: {len(tsubs)} of {b.dvm.rank} subscripts')
        rules = ', '.join(self.linear_rule(x, dummies, d.line) for x in tsubs)
        return self.line_call(d.line) + '\n' + ind + \
            f'dvmh_realign_C({a.name}, {1 if m.group(5) else 0}, {b.name}, {b.dvm.rank}, {rules});'

    def redistribute(self, d, ind):
        """redistribute(A[format]...) (cdvmh/src/converter_main.cpp:72-92)."""
        m = re.match(r'\(\s*([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)+)\)\s*$', d.rest, re.S)
        if not m or not self.scope.get(m.group(1)) or not self.scope[m.group(1)].dvm:
            self.err(d.line, 'redistribute(A[format]...) of a distributed array expected')
        a = self.scope[m.group(1)].dvm
        fm = self.dist_formats(brackets(m.group(2)), a.rank, d.line)
        return self.line_call(d.line) + '\n' + ind + f'dvmh_redistribute_C({a.var.name}, {a.rank}, {self.dist_args(fm, d.line)});'

    def rma_items(self, text, line):
        """remote_access(A[e]..., ...) items -> [(array, subscripts)]."""
        out = []
        for item in split_top(text):
            im = re.match(r'\s*([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)*)$', item)
            v = self.scope.get(im.group(1)) if im else None
            if not v or not v.dvm:
                self.err(line, f'remote_access of {item.strip()!r}, which is not a distributed array')
            subs = brackets(im.group(2))
            if subs and len(subs) != v.dvm.rank:
                self.err(line, f'{v.name}: {len(subs)} of {v.dvm.rank} subscripts')
            out.append((v, subs or [''] * v.dvm.rank))
        return out

    def remote_block(self, d, i, j, ctx):
        """A remote_access directive on a host block (converter_main.cpp:1250-1300):
        the referenced elements are read on the host, so they are made current
        there first (one process holds every element)."""
        m = re.match(r'\((.*)\)\s*$', d.rest, re.S)
        if not m or self.toks[i].text != '{':
            self.err(d.line, 'remote_access(A[...]...) followed by a compound statement expected')
        calls = [self.line_call(d.line)]
        for v, subs in self.rma_items(m.group(1), d.line):
            args = []
            for x in subs:
                x = x.strip()
                args += [f'(DvmType)({x})', f'(DvmType)({x})'] if x else ['UNDEF_BOUND', 'UNDEF_BOUND']
            calls.append(f'dvmh_get_actual_subarray_C({v.name}, {v.dvm.rank}, {", ".join(args)});')
        ind = self.indent_of(i)
        return ('\n' + ind).join(calls) + '\n' + ind + self.plain_statement(i, j, ctx)

    def actual(self, d, ind):
        fn_scalar = 'dvmh_actual_variable2_' if d.kind == 'actual' else 'dvmh_get_actual_variable2_'
        fn_array = 'dvmh_actual_array2_' if d.kind == 'actual' else 'dvmh_get_actual_array2_'
        fn_sub = 'dvmh_actual_subarray_C' if d.kind == 'actual' else 'dvmh_get_actual_subarray_C'
        m = re.match(r'\((.*)\)\s*$', d.rest, re.S)
        if not m:
            self.err(d.line, f'{d.kind} needs a parenthesised list')
        calls = [self.line_call(d.line)]
        for item in split_top(m.group(1)):
            im = re.match(r'([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)*)$', item)
            if not im or im.group(1) not in self.scope:
                self.err(d.line, f'unknown variable {item!r}')
            v = self.scope[im.group(1)]
            subs = brackets(im.group(2))
            if v.dvm:
                if not subs:
                    calls.append(f'{fn_array}({v.name});')
                else:
                    if len(subs) != v.dvm.rank:
                        self.err(d.line, f'{v.name}: {len(subs)} of {v.dvm.rank} sub-array bounds')
                    args = []
                    for s in subs:
                        lo, _, hi = s.partition(':')
                        if not _:
                            hi = lo
                        args += [f'(DvmType)({lo.strip()})' if lo.strip() else 'UNDEF_BOUND',
                                 f'(DvmType)({hi.strip()})' if hi.strip() else 'UNDEF_BOUND']
                    calls.append(f'{fn_sub}({v.name}, {v.dvm.rank}, {", ".join(args)});')
            else:
                if subs:
                    self.err(d.line, f'{v.name}: sub-ranges of non-distributed variables are not supported')
                calls.append(f'{fn_scalar}({"" if v.dims or v.pointer_rank else "&"}{v.name});')
        return ('\n' + ind).join(calls)

    def region(self, d, i, j, ctx):
        if self.toks[i].text != '{':
            self.err(d.line, 'a region directive is followed by a compound statement')
        if ctx['region'] is not None:
            self.err(d.line, 'nested regions')
        clauses = parse_clauses(d.rest)
        intents, flags, parts = {}, [], []
        for name, args in clauses:
            if name in ('in', 'out', 'inout', 'local'):
                for a in split_top(args):
                    am = re.match(r'\s*([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)+)$', a)
                    if am:
                        # a sub-array: [e] one index, [lo : hi] a range, [] the whole axis
                        parts.append((am.group(1), 'INTENT_' + name.upper(), brackets(am.group(2))))
                        continue
                    intents[a.strip()] = 'INTENT_' + name.upper()
            elif name == 'async':
                flags.append('REGION_ASYNC')
            elif name == 'targets':
                pass
            else:
                self.err(d.line, f'region clause {name!r} is not supported')
        self.region_no += 1
        rname, lname = f'dvmh_region_{self.region_no}', f'dvmh_loop_{self.region_no}'
        close = matching(self.toks, i)
        inner_ctx = {'region': rname, 'loop': lname, 'scalars': {}, 'arrays': {}}
        inner = self.block(i + 1, close, inner_ctx)
        ind = self.indent_of(i)
        ind2 = ind + '    '
        reg = [self.line_call(d.line),
               f'DvmType {rname} = dvmh_region_create_C({" | ".join(flags) if flags else "0"});',
               f'DvmType {lname} = 0;']
        arrays = dict(inner_ctx['arrays'])
        for a in list(intents) + [p[0] for p in parts]:
            if a not in self.scope or not self.scope[a].dvm:
                self.err(d.line, f'region clause names {a!r}, which is not a distributed array')
        part_names = {p[0] for p in parts}
        names = list(dict.fromkeys(list(intents) + [a for a in arrays if a not in part_names]))
        for a in names:
            intent = intents.get(a) or ('INTENT_INOUT' if arrays[a] else 'INTENT_IN')
            reg.append(f'dvmh_region_register_array_C({rname}, {intent}, {a}, "{a}");')
        for a, intent, subs in parts:
            rank = self.scope[a].dvm.rank
            if len(subs) != rank:
                self.err(d.line, f'{a}: {len(subs)} of {rank} sub-array bounds')
            bounds = []
            for x in subs:
                lo, sep, hi = x.partition(':')
                lo, hi = lo.strip(), (hi.strip() if sep else lo.strip())
                bounds += [f'(DvmType)({lo})' if lo else 'UNDEF_BOUND', f'(DvmType)({hi})' if hi else 'UNDEF_BOUND']
            reg.append(f'dvmh_region_register_subarray_C({rname}, {intent}, {a}, "{a}", {rank}, {", ".join(bounds)});')
        outs = inner_ctx.get('scalars_out', set())
        for s, ct in inner_ctx['scalars'].items():
            reg.append(f'dvmh_region_register_scalar_C({rname}, {"INTENT_INOUT" if s in outs else "INTENT_IN"}, &{s}, sizeof({s}), "{s}");')
        reg.append(f'dvmh_region_execute_on_targets_C({rname}, DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA);')
        (void_loop) = f'(void){lname};'
        return '{\n' + ''.join(ind2 + r + '\n' for r in reg) + ind2 + void_loop + inner + '\n' + ind2 + \
            self.line_call(self.toks[close].line) + '\n' + ind2 + f'dvmh_region_end_C({rname});\n' + ind + '}'

    def parallel(self, d, i, j, ctx):
        m = re.match(r'\(\s*((?:\[[^\]]*\]\s*)+)\s*on\s+([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)+)\)\s*(.*)$', d.rest, re.S)
        if not m:
            self.err(d.line, 'parallel([v]... on A[...]) expected')
        lvars = brackets(m.group(1))
        target = m.group(2)
        tsubs = brackets(m.group(3))
        clauses = parse_clauses(m.group(4))
        rank = len(lvars)
        if target not in self.scope or not self.scope[target].dvm:
            self.err(d.line, f'loop target {target!r} is not a distributed array')
        # the loop nest
        heads, k = [], i
        for q in range(rank):
            if self.toks[k].text != 'for':
                self.err(self.toks[k], f'a perfect nest of {rank} for loops expected')
            c = matching(self.toks, k + 1)
            heads.append(self.for_head(self.text(k + 2, c), lvars[q], self.toks[k].line))
            k = c + 1
            while self.toks[k].text == '{' and q < rank - 1:
                inner_close = matching(self.toks, k)
                if self.statement_end(k + 1, inner_close) != inner_close:
                    self.err(self.toks[k], 'the loop nest is not perfect')
                k += 1
        body_start = k
        body_end = self.statement_end(k, j)
        first, last = (body_start + 1, body_end - 1) if self.toks[body_start].text == '{' else (body_start, body_end)
        if first >= last:
            self.err(d.line, 'empty loop body')
        # from the start of the first body line, so every line keeps its indentation
        line0 = self.src.rfind('\n', 0, self.toks[first].start) + 1
        body = self.dedent(self.src[line0:self.toks[last - 1].end])
        reds, privates, across, shadows, cuda_block, rmas = [], [], None, [], None, []
        red_lengths = {}
        for name, args in clauses:
            if name == 'reduction':
                for r in split_top(args):
                    rm = re.match(r'(\w+)\s*\((.*)\)$', r.strip(), re.S)
                    if not rm or rm.group(1) not in RED_FUNC or rm.group(1) == 'prod':
                        self.err(d.line, f'reduction {r!r} is not supported')
                    vs = [x.strip() for x in split_top(rm.group(2))]
                    for x in vs:
                        v = self.scope.get(x)
                        if v is None or v.pointer_rank or v.dvm:
                            self.err(d.line, f'reduction variable {x!r} must be a scalar or a C array')
                        if v.dims:
                            # a reduction array: each element reduced on its own
                            n = self.const_extent(v.dims)
                            if n is None or rm.group(1) in ('maxloc', 'minloc'):
                                self.err(d.line, f'reduction array {x!r} needs one constant extent and a value reduction')
                            red_lengths[x] = n
                    if rm.group(1) in ('maxloc', 'minloc'):
                        if len(vs) != 2:
                            self.err(d.line, f'{rm.group(1)} takes a value and a loc variable')
                        reds.append((rm.group(1), vs[0], self.scope[vs[0]].ctype, vs[1], self.scope[vs[1]].ctype))
                    else:
                        if len(vs) != 1:
                            self.err(d.line, f'{rm.group(1)} takes one variable')
                        # the handler emitter's name for PRODUCT is 'prod'
                        fn = 'prod' if rm.group(1) == 'product' else rm.group(1)
                        reds.append((fn, vs[0], self.scope[vs[0]].ctype))
            elif name == 'private':
                for x in split_top(args):
                    x = x.strip()
                    if x not in self.scope:
                        self.err(d.line, f'private variable {x!r} is not declared')
                    privates.append(f'{self.scope[x].ctype} {x}')
            elif name == 'across':
                am = re.match(r'([A-Za-z_]\w*)\s*((?:\[[^\]]*\]\s*)+)$', args.strip())
                if not am:
                    self.err(d.line, 'across(A[lo:hi]...) expected')
                widths = []
                for w in brackets(am.group(2)):
                    lo, _, hi = w.partition(':')
                    widths.append((int(lo), int(hi)))
                across = {'array': am.group(1), 'widths': widths}
            elif name == 'shadow_renew':
                for x in split_top(args):
                    # A(corner): diagonal neighbours too; with one process every
                    # shadow element is already the array's own
                    x = re.sub(r'\(\s*corner\s*\)\s*$', '', x.strip()).strip()
                    if x not in self.scope or not self.scope[x].dvm:
                        self.err(d.line, f'shadow_renew of {x!r}, which is not a distributed array')
                    shadows.append(x)
            elif name == 'cuda_block':
                cuda_block = [x.strip() for x in split_top(args)]
            elif name == 'remote_access':
                rmas += self.rma_items(args, d.line)
            else:
                self.err(d.line, f'parallel clause {name!r} is not supported')
        # parameters: distributed arrays and scalars the body uses
        local = set(lvars) | {r[1] for r in reds} | {r[3] for r in reds if len(r) > 3} | \
            {p.split()[-1] for p in privates}
        local |= self.body_locals(body)
        params, written = [], set()
        inreg = ctx['region'] is not None
        for t in tokenize(body):
            if t.kind != 'id' or t.text in local or t.text not in self.scope:
                continue
            v = self.scope[t.text]
            if any(p[1] == t.text for p in params):
                continue
            if v.dvm:
                params.append(('array', t.text, v.ctype, v.dvm.rank))
            elif not v.dims and not v.pointer_rank:
                params.append(('scalar', t.text, v.ctype))
            elif not inreg and v.dims and not v.pointer_rank and not v.is_param and all(v.dims):
                # a loop outside regions runs on the host, where the program's
                # own C arrays are at hand: passed by address
                params.append(('carray', t.text, v.ctype, list(v.dims)))
            else:
                self.err(d.line, f'{t.text}: C arrays and pointers inside parallel loops are not supported'
                         + (' (inside regions)' if inreg and v.dims else ''))
        for t_ in re.finditer(r'\b([A-Za-z_]\w*)\s*(?:\[[^;]*?\]\s*)+(?:=(?!=)|\+=|-=|\*=|/=|\+\+|--)', body):
            written.add(t_.group(1))
        if target not in [p[1] for p in params]:
            params.insert(0, ('array', target, self.scope[target].ctype, self.scope[target].dvm.rank))
        # loop description for the handler emitter
        name = f'loop_{self.prog}_{d.line}'
        desc = {'name': name, 'vars': [(v, self.scope[v].ctype if v in self.scope else 'int') for v in lvars],
                'steps': [h[2] for h in heads], 'params': params, 'body': '\n' + body + '\n'}
        if reds:
            desc['reductions'] = reds
        if red_lengths:
            desc['reduction_lengths'] = red_lengths
        if across:
            desc['across'] = across
        if privates:
            desc['privates'] = privates
        if shadows:
            desc['stencil'] = shadows
        if not inreg:
            desc['host_only'] = True
        self.loops.append(desc)
        self.handlers.append((name, params, not inreg))
        # region bookkeeping
        if inreg:
            for p in params:
                if p[0] == 'array':
                    ctx['arrays'][p[1]] = ctx['arrays'].get(p[1], False) or p[1] in written
                else:
                    ctx['scalars'][p[1]] = p[2]
        # the converted statement
        ind = self.indent_of(i)
        rg = ctx['region'] if inreg else '0'
        lp = ctx['loop'] if inreg else 'dvmh_loop_0'
        out = [self.line_call(d.line)]
        for s in shadows:
            out.append(f'dvmh_shadow_renew_C({s}, 0, 0);')
        bounds = ''.join(f', (DvmType)({h[0]}), (DvmType)({h[1]}), (DvmType)({h[2]})' for h in heads)
        out.append(f'{lp} = dvmh_loop_create_C({rg}, {rank}{bounds});')
        rules = [self.linear_rule(s, lvars, d.line) for s in tsubs]
        out.append(f'dvmh_loop_map_C({lp}, {target}, {len(tsubs)}, {", ".join(rules)});')
        if across:
            ws = ''.join(f', (DvmType){lo}, (DvmType){hi}' for lo, hi in across['widths'])
            out.append(f'dvmh_loop_across_C({lp}, 0, {across["array"]}, {len(across["widths"])}{ws});')
        for v, subs in rmas:
            # with one process every element is local: the runtime only checks
            rr = ', '.join(self.linear_rule(x, lvars, d.line) for x in subs)
            out.append(f'dvmh_loop_remote_access_C({lp}, {v.name}, {v.dvm.rank}, {rr});')
        for r in reds:
            if r[2] not in RT_TYPE:
                self.err(d.line, f'reduction type {r[2]!r} is not supported')
            if len(r) > 3:
                out.append(f'dvmh_loop_reduction_C({lp}, {RED_FUNC[r[0]]}, &{r[1]}, {RT_TYPE[r[2]]}, 1, &{r[3]}, sizeof({r[3]}));')
            elif r[1] in red_lengths:
                out.append(f'dvmh_loop_reduction_C({lp}, {RED_FUNC[r[0]]}, {r[1]}, {RT_TYPE[r[2]]}, {red_lengths[r[1]]}, 0, 0);')
            else:
                out.append(f'dvmh_loop_reduction_C({lp}, {RED_FUNC[r[0]]}, &{r[1]}, {RT_TYPE[r[2]]}, 1, 0, 0);')
        if cuda_block:
            cb = (cuda_block + ['1', '1'])[:3]
            out.append(f'dvmh_loop_set_cuda_block_C({lp}, {", ".join(cb)});')
        pargs = ', '.join(p[1] if p[0] in ('array', 'carray') else '&' + p[1] for p in params)
        out.append(f'dvmh_loop_register_handler_C({lp}, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, '
                   f'dvmh_handler_func_C((DvmHandlerFunc){name}_host, {len(params)}, {pargs}));')
        if inreg:
            out.append(f'dvmh_loop_register_handler_C({lp}, DEVICE_TYPE_CUDA, 0, '
                       f'dvmh_handler_func_C((DvmHandlerFunc){name}_cuda, {len(params)}, {pargs}));')
        out.append(f'dvmh_loop_perform_C({lp});')
        if not inreg:
            return '{\n' + ind + '    DvmType dvmh_loop_0;\n' + ''.join(ind + '    ' + s + '\n' for s in out) + ind + '}'
        return ('\n' + ind).join(out)

    def for_head(self, text, var, line):
        """'i = a; i < b; i++' -> (start, end, step) texts for loop variable var."""
        parts = split_top(text, ';')
        if len(parts) != 3:
            self.err(line, f'for ({text}) is not of the form init; test; step')
        init, test, inc = parts
        mi = re.match(r'^(?:[A-Za-z_][\w\s]*\s+)?' + var + r'\s*=\s*(.+)$', init, re.S)
        if not mi:
            self.err(line, f'for loop of {var} must start with {var} = ...')
        start = mi.group(1).strip()
        mt = re.match(r'^' + var + r'\s*(<=|<|>=|>)\s*(.+)$', test, re.S)
        if not mt:
            self.err(line, f'for loop test must compare {var}')
        op, lim = mt.group(1), mt.group(2).strip()
        inc = inc.strip()
        if inc in (var + '++', '++' + var):
            step = '1'
        elif inc in (var + '--', '--' + var):
            step = '-1'
        else:
            ms = re.match(r'^' + var + r'\s*(\+=|-=)\s*(\w+)$', inc)
            if not ms:
                self.err(line, f'for loop step of {var} must be ++, --, += c or -= c')
            step = ms.group(2) if ms.group(1) == '+=' else '-' + ms.group(2)
        up = not step.startswith('-')
        if up and op in ('>', '>='):
            self.err(line, 'an increasing loop must test with < or <=')
        if not up and op in ('<', '<='):
            self.err(line, 'a decreasing loop must test with > or >=')
        end = {'<': f'({lim}) - 1', '<=': lim, '>': f'({lim}) + 1', '>=': lim}[op]
        return (start, end, step)

    def body_locals(self, body):
        """Names declared inside a loop body."""
        out = set()
        for stmt in re.split(r'[;{}]', body):
            d = parse_declaration(stmt.strip()) if stmt.strip() else None
            if d:
                out |= {x[0] for x in d}
        return out

    @staticmethod
    def dedent(text):
        lines = text.strip('\n').split('\n')
        ws = [len(re.match(r'[ \t]*', l).group()) for l in lines if l.strip()]
        cut = min(ws) if ws else 0
        return '\n'.join(l[cut:] for l in lines)

    # -- outputs
    def program_text(self, converted):
        protos = []
        for name, params, host_only in self.handlers:
            ps = ', '.join(['DvmType *'] + ['DvmType[]' if p[0] == 'array' else ('void *' if p[0] == 'carray' else f'{p[2]} *')
                                            for p in params])
            protos.append(f'void {name}_host({ps});')
            if not host_only:
                protos.append(f'void {name}_cuda({ps});')
        head = (f'/* {self.file} converted by emit/cdvmh.py: directives replaced by calls of the DVMH C\n'
                f' * interface; loop handlers are emitted from {self.prog}.loops. */\n'
                '#include "dvmhlib2.h"\n' + '\n'.join(protos) + '\n')
        if self.global_arrays:
            head += f'static void {self.prog}_globals_init(void);\n'
        return head + converted

    def loops_text(self):
        inc = [re.match(r'#\s*include\s*"([^"]+)"', x).group(1) for x in self.includes
               if re.match(r'#\s*include\s*"', x)]
        spec = {'program': self.prog, 'include': inc[0] if len(inc) == 1 else inc, 'loops': self.loops}
        if self.defines:
            spec['defines'] = self.defines
        import pprint
        return f'# Loop descriptions of {self.file}, written by emit/cdvmh.py.\n' + pprint.pformat(spec, width=110, sort_dicts=False) + '\n'


def convert_file(path, outdir, prog=None):
    """Writes <outdir>/<prog>.c and <outdir>/<prog>.loops; returns the Converter."""
    with open(path) as f:
        src = f.read()
    stem = os.path.splitext(os.path.basename(path))[0]
    prog = prog or stem
    cv = Converter(src, os.path.basename(path), prog)
    converted = cv.convert()
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, f'{prog}.c'), 'w') as f:
        f.write(cv.program_text(converted))
    with open(os.path.join(outdir, f'{prog}.loops'), 'w') as f:
        f.write(cv.loops_text())
    return cv


if __name__ == '__main__':
    if len(sys.argv) not in (3, 4):
        sys.exit('usage: cdvmh.py <program.cdv> <outdir> [program name]')
    try:
        convert_file(sys.argv[1], sys.argv[2], sys.argv[3] if len(sys.argv) == 4 else None)
    except ConvertError as e:
        sys.exit(f'cdvmh.py: {e}')
