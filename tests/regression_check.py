"""The comparison steps of the reference's regression scripts, restated for the files
this package writes (tutorials/regressionTests/singleCell/runRegressionTest.sh:31-116 and
NiedererEtAl2012/runRegressionTest.sh:29-110): for each golden row, the data row whose
time is nearest the golden time (within 1e-9 s for the trace, 2.5e-3 s for probes; '#'
lines and the trace's header line skipped) is read at the golden column and passes when
|actual - expected| <= tolerance.  Returns (passed, failed) lists of
(file, time, column, actual, expected, tolerance)."""
import os


def _rows(path, skip_first, col):
    out = []
    with open(path) as f:
        for ln, line in enumerate(f):
            if skip_first and ln == 0:
                continue
            fields = line.split()
            if not fields or fields[0].startswith("#") or len(fields) < col:
                continue
            out.append((float(fields[0]), float(fields[col - 1])))
    return out


def _check(post_dir, rows, skip_first, window):
    passed, failed = [], []
    for file_name, time, col, expected, tol in rows:
        path = os.path.join(post_dir, file_name)
        if not os.path.exists(path):
            failed.append((file_name, time, col, None, expected, tol))
            continue
        best, actual = None, None
        for t, v in _rows(path, skip_first, col):
            d = abs(t - time)
            if best is None or d < best:
                best, actual = d, v
        if best is None or best > window:
            failed.append((file_name, time, col, None, expected, tol))
        elif abs(actual - expected) <= tol:
            passed.append((file_name, time, col, actual, expected, tol))
        else:
            failed.append((file_name, time, col, actual, expected, tol))
    return passed, failed


def check_single_cell(post_dir, rows):
    """rows: (file, time, expected, tolerance); the value is column 2 (Vm)."""
    return _check(post_dir, [(f, t, 2, e, tol) for f, t, e, tol in rows], True, 1e-9)


def check_probes(post_dir, rows):
    """rows: (file, time, column, expected, tolerance); column 2 is probe 0."""
    return _check(post_dir, rows, False, 2.5e-3)
