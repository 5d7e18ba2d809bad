"""Per-launch device-time breakdown of one multigrid (or learned) flexible-CG
solve of the 128^3 pressure system, from HIP events.

A kernel trace of the flexible-CG path is not available (the profiler's host
library has faulted on it), so this tool measures the breakdown another way:
1. it copies the package's sources into a scratch directory;
2. it wraps every kernel launch of the solve's functions as
   (launch, event record);
3. it builds that copy and runs the solve once with kernel-by-kernel launches,
   after warm-up solves.
The events add about 2-3 us per launch, so the per-launch figures are upper
bounds. The solve's own time without them is printed beside them.

    python -m chuanflow_amd.launch_timeline [--precond mg|learned] [--n 128]

Measurement only: the package's own library is not touched.
"""
from __future__ import annotations

import argparse
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)

# the event recorder compiled into the scratch copy
RECORDER = r'''
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>
namespace cfx {
struct LaunchTimeline {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<hipEvent_t> ev;
    std::vector<std::string> label;
    int used = 0;
    static LaunchTimeline& get() { static LaunchTimeline t; return t; }
    void start(hipStream_t st) {
        on = std::getenv("CFX_LAUNCH_TIMELINE") != nullptr;
        if (!on || used > 0) return;
        s = st;
        mark("start");
    }
    void mark(const std::string& l) {
        if (!on) return;
        if (used >= static_cast<int>(ev.size())) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            ev.push_back(e);
            label.push_back("");
        }
        (void)hipEventRecord(ev[used], s);
        label[used] = l;
        ++used;
    }
    void report() {
        if (!on || used < 2) return;
        (void)hipStreamSynchronize(s);
        std::map<std::string, std::pair<int, double>> agg;
        std::vector<std::string> order;
        double tot = 0;
        for (int i = 1; i < used; ++i) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev[i - 1], ev[i]);
            if (!agg.count(label[i])) order.push_back(label[i]);
            auto& a = agg[label[i]];
            a.first++;
            a.second += ms;
            tot += ms;
        }
        std::printf("TIMELINE total %d %.4f\n", used - 1, tot);
        for (auto& l : order)
            std::printf("TIMELINE row %s|%d|%.3f|%.4f\n", l.c_str(), agg[l].first, agg[l].second * 1e3 / agg[l].first,
                        agg[l].second);
        std::fflush(stdout);
        used = 0;
    }
};
}
#define CFX_TL_MARK(x) ::cfx::LaunchTimeline::get().mark(x)
'''


def _func_ranges(s, names):
    out = []
    for nm in names:
        for m in re.finditer(r'\n[^\n]*\b' + re.escape(nm) + r'\s*\(', s):
            i = m.end()
            j, k = s.find('{', i), s.find(';', i)
            if j < 0 or (0 <= k < j and s[i:k].count('(') <= s[i:k].count(')')):
                continue  # a declaration or a call
            depth, t = 0, j
            while True:
                if s[t] == '{':
                    depth += 1
                elif s[t] == '}':
                    depth -= 1
                    if depth == 0:
                        break
                t += 1
            out.append((j, t))
    return out


def _close(s, q):
    # index of the parenthesis that closes the one at s[q]
    depth, t = 0, q
    while True:
        if s[t] == '(':
            depth += 1
        elif s[t] == ')':
            depth -= 1
            if depth == 0:
                return t
        t += 1


def _wrap(s, names, level):
    """(launch, mark) for every `kernel<<<...>>>(...)` and `fa.launch(...)` in the named functions."""
    edits = []
    for a, b in _func_ranges(s, names):
        body = s[a:b]
        for m in re.finditer(r'(\b[A-Za-z_][A-Za-z0-9_]*(?:<[^<>;]*>)?)<<<', body):
            st = a + m.start()
            name = re.sub(r'<.*', '', m.group(1))
            en = _close(s, s.find('>>>', st) + 3)
            lab = ('"%s L" + std::to_string(l)' % name) if level else '"%s"' % name
            edits.append((st, en + 1, lab))
        for m in re.finditer(r'\bfa\.launch\(', body):
            st = a + m.start()
            edits.append((st, _close(s, st + len('fa.launch')) + 1, '"k_fcg_A"'))
    for st, en, lab in sorted(edits, reverse=True):
        s = s[:st] + '(' + s[st:en] + ', CFX_TL_MARK(' + lab + '))' + s[en:]
    return s, len(edits)


def _replace(s, a, b):
    if a not in s:
        raise RuntimeError('launch_timeline: the source no longer holds ' + repr(a[:60]))
    return s.replace(a, b, 1)


def build_copy(dst):
    for d in ('chuanflow_amd', 'include'):
        shutil.copytree(os.path.join(REPO, d), os.path.join(dst, d),
                        ignore=shutil.ignore_patterns('_build', '_lib', '*.so', '__pycache__'))
    for f in ('bench.py', '__graft_entry__.py'):
        shutil.copy(os.path.join(REPO, f), dst)
    os.makedirs(os.path.join(dst, 'oracle'), exist_ok=True)
    dev = os.path.join(dst, 'chuanflow_amd', 'csrc', 'dev')
    with open(os.path.join(dev, 'common.h'), 'a') as f:
        f.write(RECORDER)
    counts = {}
    for fn, names, level in (('mg.hip', ['Multigrid::visit', 'Multigrid::solveLevel', 'Multigrid::launchTailUp',
                                          'Multigrid::launchUp', 'Multigrid::setWeights'], True),
                             ('mpnn.hip', ['forwardT', 'Mpnn::setGraph'], False),
                             ('pcg.hip', ['Krylov::fcgMixed'], False)):
        p = os.path.join(dev, fn)
        s = open(p).read()
        if fn == 'mg.hip':
            # setWeights has no level variable: wrap it apart
            s, n1 = _wrap(s, names[:-1], True)
            s, n2 = _wrap(s, names[-1:], False)
            counts[fn] = n1 + n2
        else:
            s, counts[fn] = _wrap(s, names, level)
        if fn == 'pcg.hip':
            s = _replace(s, '    mg.setProbe(probeFcgA);', '    ::cfx::LaunchTimeline::get().start(s);\n'
                                                            '    mg.setProbe(probeFcgA);')
            s = _replace(s, '    const bool failed = hs->nan || hs->breakdown;',
                         '    ::cfx::LaunchTimeline::get().report();\n    const bool failed = hs->nan || hs->breakdown;')
        open(p, 'w').write(s)
    p = os.path.join(dev, 'simple.hip')
    s = open(p).read()
    a = '                    a.net->setGraph(view, v, a.geo, b, sameValues);'
    s = _replace(s, a, '                    ::cfx::LaunchTimeline::get().start(m.ctx->stream());\n' + a)
    open(p, 'w').write(s)
    print('launch_timeline: building the instrumented copy', flush=True)
    r = subprocess.run([sys.executable, '-c', 'import sys; sys.path.insert(0, "."); '
                        'from chuanflow_amd.build import build_all; build_all()'], cwd=dst,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('launch_timeline: build failed\n' + r.stdout[-3000:] + r.stderr[-3000:])
    return counts


def run_copy(dst, precond, n):
    code = r'''
import os, sys, time
sys.path.insert(0, os.getcwd())
import bench as BB
from chuanflow_amd import native as N
from chuanflow_amd import learned as LN
ctx = N.Context(0)
hm, arrays, geom, dm, A, b, _ = BB.build_system(N, ctx, %d)
d_x = ctx.zeros(dm.nCells)
pc = N.PRECOND_%s
if pc == N.PRECOND_LEARNED:
    A.set_mpnn(LN.load(ctx))
r = BB.time_solve(N, ctx, A, d_x, N.FP32_IR, 3, 1, 1e-8, 20000, pc)
print("PLAIN", min(w for w, _ in r), r[-1][1]["iterations"], r[-1][1]["outerIterations"], flush=True)
os.environ["CFX_LAUNCH_TIMELINE"] = "1"
BB.time_solve(N, ctx, A, d_x, N.FP32_IR, 1, 0, 1e-8, 20000, pc)
''' % (n, 'LEARNED' if precond == 'learned' else 'MG')
    r = subprocess.run([sys.executable, '-c', code], cwd=dst, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('launch_timeline: run failed\n' + r.stdout[-3000:] + r.stderr[-3000:])
    return r.stdout


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--precond', default='learned', choices=['mg', 'learned'])
    ap.add_argument('--n', type=int, default=128)
    a = ap.parse_args(argv)
    dst = tempfile.mkdtemp(prefix='cfx_timeline_')
    try:
        counts = build_copy(dst)
        print('launch_timeline: solving', flush=True)
        out = run_copy(dst, a.precond, a.n)
    finally:
        shutil.rmtree(dst, ignore_errors=True)
    plain, rows, total = None, [], None
    for line in out.splitlines():
        if line.startswith('PLAIN'):
            _, ms, it, outer = line.split()
            plain = (float(ms), int(it), int(outer))
        elif line.startswith('TIMELINE total'):
            total = float(line.split()[3])
        elif line.startswith('TIMELINE row'):
            name, cnt, avg, tot = line[len('TIMELINE row '):].split('|')
            rows.append((name, int(cnt), float(avg), float(tot)))
    print(f'# launch timeline: {a.precond} solve at {a.n}^3 (launches wrapped: {counts})')
    print(f'solve without events: {plain[0]:.3f} ms wall, {plain[1]} FCG iterations, {plain[2]} true residuals')
    print(f'with an event after each launch: {total:.3f} ms over the launches below')
    print('| launch | count | avg us (with its event) | total ms | % |')
    print('|---|---:|---:|---:|---:|')
    for name, cnt, avg, tot in rows:
        print(f'| `{name}` | {cnt} | {avg:.2f} | {tot:.3f} | {100 * tot / total:.1f} |')
    return 0


if __name__ == '__main__':
    sys.exit(main())
