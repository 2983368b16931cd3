"""The `sdpb` command line on the GPU solver: SDPB 3.1.0's options, defaults, output files, terminate
reasons and exit codes (the reference's sdpvibe/cli.py:1-258 and __main__.py follow the same program).

    python -m sdpvibe_amd -s SDP [--precision P] [-o OUT] [-c CK] [-i INITIAL_CK]
                          [--writeSolution x,y,z,X,Y] [--maxIterations N] [--paramFile FILE] ...

Behaviour:
  * OUT defaults to SDP + "_out", CK to SDP + ".ck".
  * The run resumes from the checkpoint in INITIAL_CK (default CK) when one is there; with -i given
    explicitly a missing checkpoint is an error (exit 1).
  * A checkpoint is written every --checkpointInterval seconds (at the start of an iteration) and at
    the end unless --noFinalCheckpoint.
  * SIGTERM ends the run at the start of the next iteration; the checkpoint and the solution are
    written and the exit code is 15.
  * --procsPerNode, --procGranularity, --maxSharedMemory and --maxMemory concern SDPB's MPI processes:
    they are accepted and have no effect (the solver runs on one GPU).
  * --precision P selects L = ceil((P + 2) / 32) + 1 limbs of 32 bits (the bits asked for, two more,
    and one guard limb).
"""
import argparse
import datetime
import os
import signal
import sys
import time

import mpmath as mp

VERSION = "0.3.0"
SDPB_VERSION = "3.1.0"
SIGTERM_EXIT = 15
SIGTERM_REASON = "SIGTERM signal received"
SOLUTION_PARTS = ("x", "y", "z", "X", "Y")
_TRUE_WORDS = {"1", "true", "yes", "on"}


def _flag(text):
    return str(text).strip().lower() in _TRUE_WORDS


class Opt:
    """one sdpb option: its name, short letter, default (text, as sdpb prints it), value type"""

    def __init__(self, name, default=None, kind=str, short=None):
        self.name, self.default, self.kind, self.short = name, default, kind, short


# The option set of SDPB 3.1.0 (sdpb/SDPB_Parameters.cxx), in the order sdpb --help lists them.
OPTION_LIST = [
    Opt("sdpDir", short="s"), Opt("outDir", short="o"), Opt("paramFile", short="p"),
    Opt("checkpointDir", short="c"), Opt("initialCheckpointDir", short="i"),
    Opt("noFinalCheckpoint", False, bool), Opt("writeSolution", "x,y"), Opt("verbosity", "1"),
    Opt("procsPerNode"), Opt("procGranularity", "1"), Opt("maxSharedMemory", "0"), Opt("maxMemory", "0"),
    Opt("precision", "400", int),
    Opt("findPrimalFeasible", False, bool), Opt("findDualFeasible", False, bool),
    Opt("detectPrimalFeasibleJump", False, bool), Opt("detectDualFeasibleJump", False, bool),
    Opt("maxIterations", "500", int), Opt("maxRuntime", "9223372036854775807", int),
    Opt("checkpointInterval", "3600", int),
    Opt("dualityGapThreshold", "1e-30"), Opt("primalErrorThreshold", "1e-30"), Opt("dualErrorThreshold", "1e-30"),
    Opt("initialMatrixScalePrimal", "1e+20"), Opt("initialMatrixScaleDual", "1e+20"),
    Opt("feasibleCenteringParameter", "0.1"), Opt("infeasibleCenteringParameter", "0.3"),
    Opt("stepLengthReduction", "0.7"), Opt("minPrimalStep", "0"), Opt("minDualStep", "0"),
    Opt("maxComplementarity", "1e+100"),
]
OPTIONS = {o.name: o for o in OPTION_LIST}
VERBOSITY_LEVELS = {"0": 0, "none": 0, "1": 1, "regular": 1, "2": 2, "debug": 2, "3": 3, "trace": 3}
# the parameter listing sdpb prints before it starts (SDPB_Parameters.cxx operator<<)
PRINTED = ["maxIterations", "maxRuntime", "checkpointInterval", "maxSharedMemory", "findPrimalFeasible",
           "findDualFeasible", "detectPrimalFeasibleJump", "detectDualFeasibleJump", "precision",
           "dualityGapThreshold", "primalErrorThreshold", "dualErrorThreshold", "initialMatrixScalePrimal",
           "initialMatrixScaleDual", "feasibleCenteringParameter", "infeasibleCenteringParameter",
           "stepLengthReduction", "maxComplementarity", "initialCheckpointDir", "checkpointDir",
           "noFinalCheckpoint", "writeSolution", "procGranularity", "verbosity"]


class UsageError(Exception):
    pass


class _Parser(argparse.ArgumentParser):
    def error(self, message):
        raise UsageError(message)


def _parser():
    ap = _Parser(prog="sdpb", add_help=False, allow_abbrev=False)
    ap.add_argument("-h", "--help", action="store_true")
    ap.add_argument("--version", action="store_true")
    for o in OPTION_LIST:
        names = ["--" + o.name] + (["-" + o.short] if o.short else [])
        if o.kind is bool:                       # --flag, --flag=true, --flag=0 ...
            ap.add_argument(*names, dest=o.name, nargs="?", const=True, type=_flag, default=None)
        else:
            ap.add_argument(*names, dest=o.name, default=None)
    return ap


def _read_param_file(path):
    """'name = value' lines (# starts a comment) -> {name: text}"""
    out = {}
    with open(path) as f:
        for raw in f:
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            name, _, value = (t.strip() for t in line.partition("="))
            if name not in OPTIONS:
                raise UsageError(f"unrecognised option '{name}' in {path}")
            out[name] = value
    return out


def parse_args(argv):
    """argv -> (parameters: every option's value, command line over parameter file over default;
    {'help', 'version'}; the set of names given on the command line)"""
    ns = vars(_parser().parse_args(list(argv)))
    flags = {"help": ns.pop("help"), "version": ns.pop("version")}
    given = {k for k, v in ns.items() if v is not None}
    from_file = _read_param_file(ns["paramFile"]) if ns["paramFile"] else {}
    params = {}
    for o in OPTION_LIST:
        if ns[o.name] is not None:
            params[o.name] = ns[o.name]
        elif o.name in from_file:
            params[o.name] = _flag(from_file[o.name]) if o.kind is bool else from_file[o.name]
        else:
            params[o.name] = o.default
    return params, flags, given


def _default_dir(sdp_dir, suffix):
    p = os.path.normpath(sdp_dir)
    if os.path.basename(p) == ".":
        p = os.path.dirname(p)
    return p + suffix


def resolve(params, given):
    """fill in the defaults that depend on other options and check values"""
    if params["sdpDir"] is None:
        raise UsageError("the option '--sdpDir' is required but missing")
    params["outDir"] = params["outDir"] or _default_dir(params["sdpDir"], "_out")
    params["checkpointDir"] = params["checkpointDir"] or _default_dir(params["sdpDir"], ".ck")
    params["requireInitialCheckpoint"] = "initialCheckpointDir" in given
    params["initialCheckpointDir"] = params["initialCheckpointDir"] or params["checkpointDir"]
    level = VERBOSITY_LEVELS.get(str(params["verbosity"]).lower())
    if level is None:
        raise UsageError(f"invalid verbosity '{params['verbosity']}'")
    params["verbosity"] = level
    parts = {t.strip() for t in str(params["writeSolution"]).split(",")} - {""}
    for t in sorted(parts - set(SOLUTION_PARTS)):
        raise UsageError(f"invalid --writeSolution element '{t}' (allowed: x, y, z, X, Y)")
    params["writeSolutionParts"] = parts
    for o in OPTION_LIST:
        if o.kind is int:
            try:
                params[o.name] = int(params[o.name])
            except ValueError:
                raise UsageError(f"the argument ('{params[o.name]}') for option '--{o.name}' is invalid")
    return params


def limbs_for(precision):
    """32-bit limbs used for --precision bits: the bits asked for plus two, and one guard limb"""
    return -(-(precision + 2) // 32) + 1


def solver_params(params):
    """the solver's keyword parameters (ipm.DEFAULTS names) from the parsed options"""
    from .ipm import DEFAULTS
    return {k: params[k] if OPTIONS[k].kind in (bool, int) else str(params[k]) for k in DEFAULTS}


def options_record(params):
    """the options stored with a checkpoint (text values)"""
    rec = {}
    for o in OPTION_LIST:
        v = params.get(o.name)
        if o.name == "paramFile" or v is None:
            continue
        rec[o.name] = ("true" if v else "false") if isinstance(v, bool) else str(v)
    return rec


class Console:
    """what sdpb prints on stdout, at the chosen verbosity"""

    def __init__(self, level):
        self.level = level

    def say(self, text="", level=1):
        if self.level >= level:
            print(text, flush=True)

    @staticmethod
    def now():
        return f"{datetime.datetime.now():%Y-%b-%d %H:%M:%S}"

    def banner(self, params):
        self.say(f"{self.now()} Start SDPB (sdpvibe_amd {VERSION}, drop-in for SDPB {SDPB_VERSION})")
        self.say(f"SDPB version: {SDPB_VERSION}")
        self.say(f'SDP directory   : "{os.path.abspath(params["sdpDir"])}"')
        self.say(f'out directory   : "{os.path.abspath(params["outDir"])}"\n')
        self.say("Parameters:")
        for name in PRINTED:
            v = params[name]
            if isinstance(v, bool):
                v = "true" if v else "false"
            elif name.endswith("heckpointDir"):
                v = f'"{os.path.abspath(v)}"'
            label = name
            if name == "precision":
                label, v = "precision(actual)", f"{v}({v})"
            self.say(f"{label:<29}= {v}")
        self.say()

    def table_head(self):
        self.say(f"{self.now()} Start solver iterations")
        self.say("\n          time    mu     P-obj       D-obj      gap         P-err       p-err       D-err      "
                 "P-step   D-step   beta\n" + "-" * 117)

    def table_row(self, row):
        v = {k: float(mp.mpf(row[k])) for k in ("mu", "P-obj", "D-obj", "gap", "P-err", "p-err", "D-err", "P-step",
                                                 "D-step", "beta")}
        cells = [f"{row['iteration']:<4} ", f"{int(row['total_time']):>8}", f"{v['mu']:<#8.2g}"]
        cells += [f"{v[k]:<+#11.3g}" for k in ("P-obj", "D-obj")] + [f"{v['gap']:<#10.3g}"]
        cells += [f"{v[k]:<+#11.3g}" for k in ("P-err", "p-err", "D-err")]
        cells += [f"{v['P-step']:<#8.3g}", f"{v['D-step']:<#8.3g}", f"{v['beta']:<#4.3g}"]
        self.say(" ".join(cells))

    def verdict(self, res, fmt):
        self.say(f"-----{res['terminateReason']}-----\n")
        for key, label in (("primalObjective", "primalObjective "), ("dualObjective", "dualObjective   "),
                           ("dualityGap", "dualityGap      "), ("primalError", "primalError     "),
                           ("dualError", "dualError       ")):
            self.say(f"{label}= {fmt(res[key])}")
        self.say()


class Run:
    """one sdpb invocation after its options are resolved: load the SDP, resume, iterate, write"""

    def __init__(self, params, start):
        self.p, self.start = params, start
        self.out = Console(params["verbosity"])
        self.L = limbs_for(params["precision"])
        self.stop_requested = False

    def _on_sigterm(self, *_):
        self.stop_requested = True

    def checkpoint(self, solver):
        from . import checkpoint as CK
        from . import state as STATE
        self.out.say(f"Saving checkpoint to    : {self.p['checkpointDir']}")
        CK.save(self.p["checkpointDir"], STATE.get_state(solver), self.sdp, self.L, self.p["precision"],
                options_record(self.p))

    def setup(self):
        """-> error text or None"""
        from . import checkpoint as CK
        from . import state as STATE
        from .ipm import Solver
        from .sdp import SDP
        p = self.p
        self.out.banner(p)
        try:
            self.sdp = SDP(p["sdpDir"], self.L)
        except (ValueError, OSError, KeyError) as e:
            return str(e)
        sdp = self.sdp
        self.out.say(f"Initialize SDP solver\n\tprimal dimension: {sum(g.nb * g.P for g in sdp.groups)}"
                     f"\n\tdual dimension: {sdp.N}\n\tSDP blocks: {sdp.nblocks}")
        if getattr(sdp, "read_info", None):
            self.out.say(f"\t{sdp.read_info}", level=2)
        if "z" in p["writeSolutionParts"] and sdp.normalization is None:
            return "Please provide SDP with valid normalization.json or exclude z from --writeSolution arguments."
        self.solver = Solver(sdp, self.L, precision=p["precision"], **solver_params(p))
        try:
            st = CK.load(p["initialCheckpointDir"], sdp, self.L)
        except (ValueError, OSError) as e:
            return f"cannot load checkpoint from {p['initialCheckpointDir']}: {e}"
        if st is None:
            if p["requireInitialCheckpoint"]:
                return f"Unable to load checkpoint from directory: {p['initialCheckpointDir']}"
        else:
            self.out.say(f"Loading checkpoint from : {st['source']}")
            STATE.set_state(self.solver, st)
        return None

    def solve(self):
        from . import sdpb_output as O
        log = O.IterationsJson(self.p["outDir"])
        self.out.table_head()

        def row_done(row):
            log.append(row)
            self.out.table_row(row)

        previous = signal.signal(signal.SIGTERM, self._on_sigterm)
        try:
            res = self.solver.run(on_iteration=row_done, max_runtime=self.p["maxRuntime"],
                                  checkpoint=self.checkpoint, checkpoint_interval=self.p["checkpointInterval"],
                                  stop=lambda: self.stop_requested)
        finally:
            signal.signal(signal.SIGTERM, previous)
            log.close()
        return res

    def write(self, res):
        from . import sdpb_output as O
        from . import state as STATE
        from .ipm import fmt_big
        p, solver = self.p, self.solver
        fmt = lambda v: fmt_big(v, solver.digits)
        self.out.verdict(res, fmt)
        stopped = res["terminateReason"] == SIGTERM_REASON
        if stopped or not p["noFinalCheckpoint"]:
            self.checkpoint(solver)
        out_dir = p["outDir"]
        self.out.say(f'Saving solution to      : "{os.path.abspath(out_dir)}"')
        os.makedirs(out_dir, exist_ok=True)
        O.write_out_txt(out_dir, res, solver.digits, time.time() - self.start)
        if p["writeSolutionParts"]:
            O.write_solution(out_dir, STATE.get_state(solver), p["writeSolutionParts"], solver.digits,
                             self.sdp.normalization)
        O.write_c_minus_By(os.path.join(out_dir, "c_minus_By", "c_minus_By.json"), STATE.c_minus_By_blocks(solver),
                           solver.digits)
        if stopped:
            self.out.say("Received SIGTERM, exiting gracefully...")
            return SIGTERM_EXIT
        return 0


def _help_text():
    lines = [__doc__, f"Options (as sdpb {SDPB_VERSION}):"]
    for o in OPTION_LIST:
        arg = "" if o.kind is bool else (f" arg (={o.default})" if o.default is not None else " arg")
        lines.append(f"  --{o.name}{arg}")
    return "\n".join(lines)


def main(argv=None):
    """sdpb's exit codes: 0 done, 1 usage or input error, 15 stopped by SIGTERM"""
    start = time.time()
    try:
        params, flags, given = parse_args(sys.argv[1:] if argv is None else argv)
        if flags["help"]:
            print(_help_text())
            return 0
        if flags["version"]:
            print(f"sdpvibe_amd {VERSION} (drop-in for SDPB {SDPB_VERSION})")
            return 0
        params = resolve(params, given)
    except (UsageError, OSError) as e:
        print(f"Error: {e}\nRun with --help for the list of options.", file=sys.stderr)
        return 1
    run = Run(params, start)
    err = run.setup()
    if err:
        print(f"Error: {err}", file=sys.stderr)
        return 1
    return run.write(run.solve())
