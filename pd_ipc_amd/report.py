"""Headless pd_bench-style report of one cloth on the GPU.

``python -m pd_ipc_amd.report [pd_bench options]`` runs the cloth pd_bench runs
(app/bench/main.cpp:53-264: the same options, defaults and iteration presets,
the scene makeScene builds) through the library's ``pdx_bench_run`` and prints
its report: setup, results (converged / total substeps, last residual, final
and run-max strain, elastic and total energy), stage times and factor reuse.
``run(argv)`` returns the report text and its figures as a dict.

Options beyond pd_bench: ``--shear K`` / ``--bend K`` (pd_bench reads PD_SHEAR /
PD_BEND from the environment; those are honoured too), ``--precision
fp64|mixed32``, ``--solve exact|fast``, ``--device D``.
"""
import ctypes as C
import sys

from .pdx import PdxError, _BenchSummary, load_library

_TEXT_CAP = 1 << 20


def run(argv):
    """Runs pdx_bench_run with the given options (without the program name)."""
    lib = load_library()
    args = [b"pdx_bench"] + [a.encode() if isinstance(a, str) else a for a in argv]
    arr = (C.c_char_p * len(args))(*args)
    buf = C.create_string_buffer(_TEXT_CAP)
    summary = _BenchSummary()
    rc = lib.pdx_bench_run(len(args), arr, buf, _TEXT_CAP, C.byref(summary))
    text = buf.value.decode(errors="replace")
    if rc != 0:
        raise PdxError(rc, text.strip())
    return text, {k: getattr(summary, k) for k, _ in _BenchSummary._fields_}


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    try:
        text, _ = run(argv)
    except PdxError as e:
        sys.stderr.write(str(e) + "\n")
        return 2
    sys.stdout.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
