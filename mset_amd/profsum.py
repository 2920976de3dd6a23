"""Text summaries of rocprofv3 result databases (the SQLite files `rocprofv3 -d DIR -o NAME` writes).

    python -m mset_amd.profsum trace DB                       per-kernel calls, total / average duration
    python -m mset_amd.profsum counters DB [--kernel REGEX]   per-dispatch counter means of matching kernels
    python -m mset_amd.profsum calibrate FETCH_DB WRITE_DB    counter / known-bytes factors of mset_calibrate

Counter values are reported as rocprofv3 stores them (FETCH_SIZE and WRITE_SIZE in KiB per dispatch).
"""
import argparse
import re
import sqlite3
import sys

# bytes each kernel of mset_amd/mset_calibrate moves when run with its default size (1024 MiB)
CALIBRATION = {'k_cal_copy': (1 << 30, 1 << 30), 'k_cal_read': (1 << 30, 4096 * 4), 'k_cal_write': (0, 1 << 30),
               'k_cal_tile': (1 << 30, 1 << 30)}


def _rows(db, sql):
    con = sqlite3.connect(db)
    try:
        return con.execute(sql).fetchall()
    finally:
        con.close()


def trace(db):
    rows = _rows(db, 'select name, duration from kernels')
    tot = {}
    for name, dur in rows:
        n, t = tot.get(name, (0, 0.0))
        tot[name] = (n + 1, t + float(dur))
    gpu = sum(t for _, t in tot.values())
    out = ['# name | calls | total ms | avg us | % of GPU time']
    for name, (n, t) in sorted(tot.items(), key=lambda kv: -kv[1][1]):
        out.append('%s | %d | %.3f | %.1f | %.2f' % (name, n, t / 1e6, t / n / 1e3, 100.0 * t / gpu))
    return out


def counters(db, kernel=None):
    rows = _rows(db, 'select k.name, e.dispatch_id, e.counter_name, e.counter_value, k.duration '
                     'from pmc_events e join kernels k on k.dispatch_id = e.dispatch_id')
    per = {}
    for name, disp, cname, val, dur in rows:
        if kernel and not re.search(kernel, name):
            continue
        per.setdefault((name, cname), {}).setdefault(disp, 0.0)
        per[(name, cname)][disp] += float(val)      # counter instances summed per dispatch
    out = ['# kernel | counter | dispatches | mean | min | max (per dispatch)']
    for (name, cname), d in sorted(per.items()):
        v = list(d.values())
        out.append('%s | %s | %d | %.6g | %.6g | %.6g' % (name, cname, len(v), sum(v) / len(v), min(v), max(v)))
    return out


def calibrate(fetch_db, write_db):
    out = ['# kernel | known read bytes | FETCH_SIZE KiB | read factor | known write bytes | WRITE_SIZE KiB | write factor']
    f = {r[0]: r for r in _cal_rows(fetch_db, 'FETCH_SIZE')}
    w = {r[0]: r for r in _cal_rows(write_db, 'WRITE_SIZE')}
    for k, (rb, wb) in CALIBRATION.items():
        fk = f.get(k, (k, float('nan')))[1]
        wk = w.get(k, (k, float('nan')))[1]
        rf = rb / (fk * 1024.0) if rb and fk == fk and fk > 0 else float('nan')
        wf = wb / (wk * 1024.0) if wb and wk == wk and wk > 0 else float('nan')
        out.append('%s | %d | %.6g | %.4f | %d | %.6g | %.4f' % (k, rb, fk, rf, wb, wk, wf))
    return out


def _cal_rows(db, counter):
    """(kernel, mean per dispatch) of a counter for the calibration kernels (instances summed per dispatch)."""
    per = {}
    for name, disp, val in _rows(db, "select k.name, e.dispatch_id, e.counter_value from pmc_events e join kernels k "
                                     "on k.dispatch_id = e.dispatch_id where e.counter_name = '%s'" % counter):
        m = re.search(r'(k_cal_\w+)', name)
        if m:
            d = per.setdefault(m.group(1), {})
            d[disp] = d.get(disp, 0.0) + float(val)
    return [(k, sum(d.values()) / len(d)) for k, d in per.items()]


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    a = sub.add_parser('trace')
    a.add_argument('db')
    a = sub.add_parser('counters')
    a.add_argument('db')
    a.add_argument('--kernel', default=None)
    a = sub.add_parser('calibrate')
    a.add_argument('fetch_db')
    a.add_argument('write_db')
    args = ap.parse_args(argv)
    if args.cmd == 'trace':
        lines = trace(args.db)
    elif args.cmd == 'counters':
        lines = counters(args.db, args.kernel)
    else:
        lines = calibrate(args.fetch_db, args.write_db)
    print('\n'.join(lines))
    return 0


if __name__ == '__main__':
    sys.exit(main())
