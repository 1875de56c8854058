"""TOML job specification and the reference's run sequence on the MI355X engine.

`load` reads a configuration file with the keys, defaults and checks of
DqmcInitializer::parse_toml_config (src/dqmc_initializer.cpp:27-338): [Model] type and
Params (hopping_t, onsite_u, chemical_potential), [Lattice] type, cell, momentum,
momentum_list, [CheckerBoard] whether_or_not, [MonteCarlo] beta, time_size,
stabilization_pace, [Measure] sweeps_warmup, bin_num, bin_size, sweeps_between_bins,
observables.  `run` is src/dqmc_main.cpp:110-360 with the MPI ranks replaced by chains of one
batch on the GPU: every chain thermalizes and measures its share of the bins, the bins of all
chains are analysed together, and chain 0's fields and the observables' mean and bin files are
written in the reference's formats.  `job` turns a specification into a serving job.

Command line (the reference's options):
  python jobspec.py --config config.toml --output out_dir [--fields fields.out] [--chains K] [--seed S]
"""
import argparse
import json
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dqmc_amd as D  # noqa: E402

try:
    import tomllib as _toml          # Python >= 3.11
except ImportError:                  # pragma: no cover - depends on the interpreter
    import tomli as _toml

EQUALTIME = ["filling_number", "double_occupancy", "kinetic_energy", "momentum_distribution", "local_spin_corr",
             "spin_density_structure_factor", "charge_density_structure_factor", "s_wave_pairing_corr"]
DYNAMIC = ["greens_functions", "density_of_states", "superfluid_stiffness", "dynamic_spin_susceptibility"]
ALL = EQUALTIME + DYNAMIC   # Measure::MeasureHandler::ObservableAll order (src/measure/observable_handler.cpp:32-45)
DESCRIPTION = {"equaltime_sign": "Averaged sign (equal-time)", "dynamic_sign": "Averaged sign (dynamical)",
               "filling_number": "Filling number", "double_occupancy": "Double occupation",
               "kinetic_energy": "Kinetic energy", "local_spin_corr": "Local spin correlation",
               "momentum_distribution": "Momentum distribution", "spin_density_structure_factor": "SDW order parameter",
               "charge_density_structure_factor": "CDW order parameter",
               "s_wave_pairing_corr": "S-wave pairing correlation", "superfluid_stiffness": "Superfluid stiffness"}
# engine index of each equal-time observable (dqmc_get_measure order)
EQ_INDEX = {"equaltime_sign": 0, "filling_number": 1, "double_occupancy": 2, "kinetic_energy": 3,
            "momentum_distribution": 4, "local_spin_corr": 5, "spin_density_structure_factor": 6,
            "charge_density_structure_factor": 7, "s_wave_pairing_corr": 8}


class ConfigError(ValueError):
    pass


def num_kstars(side):
    return (side // 2 + 1) * (side // 2 + 2) // 2


def momentum_index(side, name):
    """Lattice::Square high-symmetry points (src/lattice/square.cpp:108-114)."""
    nk = num_kstars(side)
    points = {"GammaPoint": 0, "XPoint": nk - side // 2 - 1, "MPoint": nk - 1}
    if name not in points:
        raise ConfigError(f"undefined momentum '{name}' for 2d square lattice, please check the config.")
    return points[name]


def momentum_list(side, name):
    """Lattice::Square high-symmetry lines and the loop Gamma -> X -> M -> Gamma, as k-star
    indices (src/lattice/square.cpp:116-145)."""
    h = side // 2
    x = momentum_index(side, "XPoint")
    lists = {"KstarsAll": list(range(num_kstars(side))),
             "DeltaLine": [i * (i + 1) // 2 for i in range(h + 1)],
             "ZLine": [x + i for i in range(h + 1)],
             "SigmaLine": [i * (i + 3) // 2 for i in range(h + 1)],
             "Gamma2X2M2GammaLoop": ([i * (i + 1) // 2 for i in range(h)] + [x + i for i in range(h)]
                                     + [i * (i + 3) // 2 for i in range(h, 0, -1)])}
    if name not in lists:
        raise ConfigError(f"undefined momentum list '{name}' for 2d square lattice, please check the config.")
    return lists[name]


class JobSpec:
    """The parsed configuration: chain parameters plus the measuring schedule."""

    def __init__(self, cfg, world_size=1):
        model = cfg.get("Model", {})
        mtype = model.get("type", "RepulsiveHubbard")
        if mtype not in ("RepulsiveHubbard", "AttractiveHubbard"):
            raise ConfigError(f"undefined model type '{mtype}', please check the config.")
        prm = model.get("Params", {})
        self.model = D.REPULSIVE if mtype == "RepulsiveHubbard" else D.ATTRACTIVE
        self.hopping_t = float(prm.get("hopping_t", 1.0))
        self.onsite_u = float(prm.get("onsite_u", 4.0))
        self.chemical_potential = float(prm.get("chemical_potential", 0.0))
        lat = cfg.get("Lattice", {})
        ltype = lat.get("type", "Square")
        if ltype in ("Cubic", "Honeycomb"):
            raise ConfigError(f"lattice type '{ltype}' is not supported by the MI355X engine (square lattices only).")
        if ltype != "Square":
            raise ConfigError(f"undefined lattice type '{ltype}', please check the config.")
        cell = lat.get("cell")
        if (not isinstance(cell, list) or len(cell) != 2 or not all(isinstance(c, int) and not isinstance(c, bool)
                                                                   for c in cell) or min(cell) < 1):
            raise ConfigError("the input cell of 2d square lattice should be a vector containing two positive "
                              "intergers, please check the config.")
        if cell[0] != cell[1]:
            raise ConfigError("the engine's square lattice has equal sides (cell = [L, L]).")
        self.side = cell[0]
        self.checkerboard = bool(cfg.get("CheckerBoard", {}).get("whether_or_not", False))
        mc = cfg.get("MonteCarlo", {})
        self.beta = float(mc.get("beta", 4.0))
        self.time_size = int(float(mc.get("time_size", 80)))   # read as a double, truncated (:192-194)
        self.pace = int(mc.get("stabilization_pace", 10))
        meas = cfg.get("Measure", {})
        self.sweeps_warmup = int(meas.get("sweeps_warmup", 512))
        self.bin_num = int(meas.get("bin_num", 20))
        self.bin_size = int(meas.get("bin_size", 100))
        self.sweeps_between_bins = int(meas.get("sweeps_between_bins", 20))
        obs = meas.get("observables")
        if not isinstance(obs, list) or not all(isinstance(o, str) for o in obs):
            raise ConfigError("undefined observables, please check the config.")
        if len(obs) == 1 and obs[0] in ("all", "All"):
            obs = list(ALL)
        elif len(obs) == 1 and obs[0] in ("none", "None"):
            obs = []
        unknown = [o for o in obs if o not in ALL]
        if unknown:
            raise ConfigError(f"undefined observable '{unknown[0]}', please check the config.")
        self.observables = obs
        self.bins_per_proc = math.ceil(self.bin_num / world_size)   # src/dqmc_initializer.cpp:244
        self.momentum = momentum_index(self.side, lat.get("momentum", ""))
        self.momentum_list = momentum_list(self.side, lat.get("momentum_list", ""))

    @property
    def is_equaltime(self):
        return any(o in EQUALTIME for o in self.observables)

    @property
    def is_dynamic(self):
        return any(o in DYNAMIC for o in self.observables)

    def params(self, seed=0, delay=16, precision=D.PREC_FP64):
        return D.make_params(self.side, self.time_size, self.beta, pace=self.pace, model=self.model,
                             checkerboard=self.checkerboard, delay=delay, momentum=self.momentum, t=self.hopping_t,
                             U=self.onsite_u, mu=self.chemical_potential, seed=seed, precision=precision)


def load(path=None, text=None, world_size=1):
    if text is None:
        with open(path, "rb") as f:
            cfg = _toml.load(f)
    else:
        cfg = _toml.loads(text)
    return JobSpec(cfg, world_size)


def job(spec, seed=0, precision=D.PREC_FP64):
    """A serving job (Server.submit arguments) for the specification: its warm-up, then
    bin_num x bin_size measured sweeps in bins of bin_size, with the dynamic option when the
    specification asks for dynamic observables (the server runs no sweeps between bins)."""
    return dict(params=spec.params(seed=seed, precision=precision), warmup=2 * (spec.sweeps_warmup // 2),
                measure=spec.bin_num * 2 * (spec.bin_size // 2), bin_size=2 * (spec.bin_size // 2),
                dynamic=spec.is_dynamic)


def run(spec, ctx, chains=1, seed=0, fields=None, out_dir=None, log=None):
    """Dqmc::thermalize, Dqmc::measure (bins of bin_size/2 forth-and-back iterations, the
    dynamic sweep opening each when dynamic observables are measured, sweeps_between_bins/2
    unmeasured iterations after each bin) and Dqmc::analyse over the bins of every chain.
    `fields` (L x N of +-1) starts every chain from that configuration.  Returns the means and
    errors; with out_dir also writes the reference's output files."""
    per_chain = math.ceil(spec.bin_num / chains)
    b = D.Batch(D.Context(0) if ctx is None else ctx, [spec.params(seed=seed + c) for c in range(chains)])
    if fields is None:
        b.fields_random()
    else:
        for c in range(chains):
            b.set_fields(c, fields)
    b.init()
    b.sweep(2 * (spec.sweeps_warmup // 2))
    eq_bins, dyn_bins = [], []
    if spec.is_equaltime or spec.is_dynamic:
        for _ in range(per_chain):
            b.clear_measure()
            b.sweep(2 * (spec.bin_size // 2), measure=True, dynamic=spec.is_dynamic)
            for c in range(chains):
                s, n = b.measure_accum(c)
                if spec.is_equaltime:
                    eq_bins.append(D.normalize_bin(s, n))
                if spec.is_dynamic:
                    a = b.dynamic_accum(c, spec.momentum_list)
                    dyn_bins.append(dict(sign=abs(a["sign"] / a["count"]), rho=a["rho"] / a["sign"],
                                         gk=a["gk"] / a["sign"], dos=a["dos"] / a["sign"], dss=a["dss"] / a["sign"]))
            b.sweep(2 * (spec.sweeps_between_bins // 2))
            if log:
                print(f"[jobspec] bin done ({len(eq_bins) or len(dyn_bins)} bins)", file=log, flush=True)
    res = {"bins": max(len(eq_bins), len(dyn_bins)),
           "wrap_error": max(b.info(c).wrap_error for c in range(chains))}
    if eq_bins:
        m, e = D.bin_stats(eq_bins)
        for name, k in EQ_INDEX.items():
            if name == "equaltime_sign" or name in spec.observables:
                res[name] = (float(m[k]), float(e[k]))
    if dyn_bins:
        for key, name in (("sign", "dynamic_sign"), ("rho", "superfluid_stiffness"), ("gk", "greens_functions"),
                          ("dos", "density_of_states"), ("dss", "dynamic_spin_susceptibility")):
            if key != "sign" and name not in spec.observables:
                continue
            m, e = D.bin_stats(np.array([bb[key] for bb in dyn_bins]))
            res[name] = (m, e)
            res[name + "_bins"] = np.array([bb[key] for bb in dyn_bins])
    if eq_bins and "s_wave_pairing_corr" in spec.observables:
        res["s_wave_pairing_corr_bins"] = np.array([x[EQ_INDEX["s_wave_pairing_corr"]] for x in eq_bins])
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        D.write_fields(os.path.join(out_dir, "fields.out"), b.fields(0))
        D.write_kstars(os.path.join(out_dir, "kstars.out"), spec.side)
        D.write_tgrids(os.path.join(out_dir, "tgrids.out"), spec.time_size, spec.beta)
        for name, stem in (("s_wave_pairing_corr", "swave"), ("density_of_states", "dos"), ("greens_functions", "greens"),
                           ("dynamic_spin_susceptibility", "dss")):
            if name in res:
                D.write_observable(os.path.join(out_dir, stem + ".out"), *res[name])
                D.write_bins(os.path.join(out_dir, stem + ".bins.out"), res[name + "_bins"])
    b.close()
    return res


def report(res, out=sys.stdout):
    """The scalar lines of the reference's screen output (DqmcIO::output_observable, std::ostream)."""
    for name in ["equaltime_sign", "dynamic_sign", None, "filling_number", "double_occupancy", "kinetic_energy",
                 "local_spin_corr", "momentum_distribution", "spin_density_structure_factor",
                 "charge_density_structure_factor", "s_wave_pairing_corr", "superfluid_stiffness"]:
        if name is None:
            print(file=out)
        elif name in res:
            m, e = res[name]
            m, e = float(np.ravel(m)[0]), float(np.ravel(e)[0])
            print(f"{DESCRIPTION[name]:>30s}{'->':>7s}{m:20.12f}  pm  {e:.12f}", file=out)


def main():
    ap = argparse.ArgumentParser(description="Run a dqmc-framework configuration on the MI355X engine.")
    ap.add_argument("--config", "-c", default="config.toml")
    ap.add_argument("--output", "-o", default=".")
    ap.add_argument("--fields", "-f", default=None, help="field configuration file (random fields if absent)")
    ap.add_argument("--chains", type=int, default=0, help="chains sharing the bins (default: one per bin)")
    ap.add_argument("--seed", type=int, default=0, help="seed of chain 0 (chain c uses seed + c)")
    args = ap.parse_args()
    spec = load(args.config)
    chains = args.chains or max(1, spec.bin_num)
    fields = D.read_fields(args.fields, spec.time_size, spec.side * spec.side) if args.fields else None
    ctx = D.Context(0)
    res = run(spec, ctx, chains=chains, seed=args.seed, fields=fields, out_dir=args.output, log=sys.stderr)
    report(res)
    print(json.dumps({k: [float(np.ravel(v[0])[0]), float(np.ravel(v[1])[0])] for k, v in res.items()
                      if isinstance(v, tuple)}))


if __name__ == "__main__":
    main()
