"""Molecular-system setup: YAML dataset -> species, monomers, states, bonds.

Host-side restatement of the reference front end that feeds the SCF path.  It
reproduces the index conventions the hot path depends on:

* monomer ordering = codepoint sort of every letter used in a sequence plus
  every species named in a reaction
  (src/molecular_system/setup/species/chemistry/used_monomers.jl:27-49);
* state families = connected components of the reaction graph, restricted to
  monomers that appear in sequences
  (src/molecular_system/setup/species/chemistry/state_families.jl:1-90);
* delta_muH per monomer from pK / pSpecies
  (src/molecular_system/setup/species/properties/assign_delta_muH.jl:50-247);
* topology: root = segment 1, children/parents/levels
  (src/molecular_system/setup/configurations/parse_topology.jl:16-127);
* bond types = sorted unique (i <= j) bead-type pairs
  (src/molecular_system/setup/configurations/process_bonds.jl:8-29);
* freely-jointed max_distance (BFS along bonds)
  (src/molecular_system/setup/configurations/species_model/freely_jointed.jl:9-114).

All indices here are 0-based; the reference is 1-based.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

LOG10 = math.log(10.0)

# src/molecular_system/structures/ConstantsStruct.jl:18-24
EPSILON0 = 8.854187817e-12
E0 = 1.60217653e-19
K_B = 1.3806505e-23
N_A = 6.0221415e23

KNOWN_FE_MODELS = ("ideal", "mfmt", "TPT1mfmt", "msa", "TPT1msa", "sw")


class SetupError(ValueError):
    """Raised for inputs the reference setup rejects (it calls `error()`)."""


@dataclass
class Constants:
    l_sc: float = 0.5e-9  # metres

    def concentration_to_density(self, conc: float) -> float:
        # src/molecular_system/setup/core/conversions/conversions.jl:19-22
        return conc * N_A * (self.l_sc ** 3 * 1000.0)

    def density_to_concentration(self, rho: float) -> float:
        return rho / (self.l_sc ** 3 * 1000.0) / N_A


@dataclass
class Topology:
    monomers: str
    children: List[List[int]]
    parents: List[List[int]]
    levels: List[List[int]]


@dataclass
class Species:
    name: str                       # the sequence string
    sequence: List[int]             # monomer index per segment
    state_family: List[List[int]]   # allowed bead types per segment
    topology: Topology
    model: str                      # "freely_jointed" | "monomerbead"
    max_distance: Optional[np.ndarray]
    input_density: float
    pSpecies: float                 # NaN when absent
    adjustable: bool
    role: str
    fixed: Optional[dict] = None    # {"density", "segment", "position"} from YAML
    evaluation: str = "analytical"

    @property
    def n_seg(self) -> int:
        return len(self.sequence)


@dataclass
class Reaction:
    species: List[str]
    coeffs: List[int]
    pK: float


@dataclass
class MolSys:
    constants: Constants
    temperature: float
    dielectric: float
    bjerrum: float
    monomers: List[str]             # sorted monomer names; index = bead type
    diameters: np.ndarray
    valences: np.ndarray
    delta_muH: np.ndarray
    species: List[Species]
    reactions: List[Reaction]
    bond_types: List[Tuple[int, int]]
    fe_model: List[str]
    species_charge: np.ndarray = field(default_factory=lambda: np.zeros(0))
    sw: Optional["SquareWellPairs"] = None    # pairs: square_well (process_pairs.jl:42-95)

    @property
    def NB(self) -> int:
        return len(self.monomers)

    @property
    def MB(self) -> int:
        return len(self.bond_types)

    def bond_index(self, si: int, sj: int) -> Tuple[int, int, int]:
        """(bond, pair_i, pair_j) with pair_* in {0, 1}; get_bond_info
        (src/iDFT/coordinate_system/Cartesian/species_model/freely_jointed/child_propagators.jl:204-215)."""
        if si <= sj:
            pair, pi, pj = (si, sj), 0, 1
        else:
            pair, pi, pj = (sj, si), 1, 0
        return self.bond_types.index(pair), pi, pj


@dataclass
class SquareWellPairs:
    """properties.pairs[:square_well] (src/molecular_system/setup/species/properties/process_pairs.jl:42-95):
    ordered bead-type pairs (i <= j, 0-based), their well depths and width multipliers, and the
    bead types that take part."""
    pairs: List[Tuple[int, int]]
    energys: List[float]
    lambdas: List[float]
    species: List[int]

    def index(self, i: int, j: int) -> Optional[int]:
        """findfirst(==((min, max)), pairs): the first matching pair, None if absent."""
        key = (min(i, j), max(i, j))
        for k, pr in enumerate(self.pairs):
            if pr == key:
                return k
        return None


def process_square_well(monomers: List[str], dataset: dict) -> Optional[SquareWellPairs]:
    """process_pairs.jl:42-95 for the `square_well` pair type. Pairs are visited as Julia's
    Iterators.product(sorted monomers, sorted monomers) visits them (first name fastest), a YAML key
    "A, B" naming an ordered pair; each interaction's properties are appended to per-property lists
    (energy -> energys, lambda -> lambdas); lambdas default to 1.5 when no pair gives one."""
    pair_data = dataset.get("pairs") or {}
    if "square_well" not in pair_data:
        return None
    pd = pair_data["square_well"] or {}
    pairs: List[Tuple[int, int]] = []
    props: Dict[str, List[float]] = {}
    for j2, name2 in enumerate(monomers):
        for j1, name1 in enumerate(monomers):
            key = f"{name1}, {name2}"
            if key not in pd:
                continue
            pairs.append((j1, j2) if j1 < j2 else (j2, j1))
            for prop, value in (pd[key] or {}).items():
                props.setdefault(str(prop) + "s", []).append(float(value))
    energys = props.get("energys", [])
    if "energys" in props and "lambdas" not in props:
        props["lambdas"] = [1.5] * len(energys)
    lambdas = props.get("lambdas", [])
    if len(energys) < len(pairs) or len(lambdas) < len(pairs):
        raise SetupError("square_well: every pair needs an energy (and a lambda once any pair gives one)")
    involved = sorted({i for pr in pairs for i in pr})
    return SquareWellPairs(pairs=pairs, energys=energys, lambdas=lambdas, species=involved)


# ---------------------------------------------------------------- topology

def parse_topology(sequence: str) -> Topology:
    """Tree topology of a sequence string such as "AAAB", "A(BC)2D".

    Restates parse_topology (src/molecular_system/setup/configurations/parse_topology.jl:16-127),
    including its bookkeeping quirk that a trailing segment appends an extra
    empty children list (harmless: children is indexed by segment only)."""
    monomers: List[str] = []
    children: List[List[int]] = []
    parents: List[List[int]] = []
    levels: List[List[int]] = []
    counter = [0]

    def matching(s: str) -> int:
        bal = 1
        for j in range(1, len(s)):
            if s[j] == "(":
                bal += 1
            elif s[j] == ")":
                bal -= 1
            if bal == 0:
                return j
        raise SetupError(f"Unmatched parenthesis in: {s}")

    def rec(sub: str, level: int, parent: int) -> None:
        ch = sub[0]
        if ch.isalpha():
            counter[0] += 1
            cur = counter[0]
            monomers.append(ch)
            if level > len(levels):
                levels.append([])
            levels[level - 1].append(cur)
            if parent == 0:
                children.append([])
                parents.append([])
            else:
                if parent > len(children):
                    children.append([cur])
                else:
                    children[parent - 1].append(cur)
                if cur > len(parents):
                    parents.append([parent])
                else:
                    parents[cur - 1].append(parent)
            if len(sub) > 1:
                rec(sub[1:], level + 1, cur)
            else:
                children.append([])
        elif ch == "(":
            end = matching(sub)
            group = sub[1:end]
            rest = sub[end + 1:]
            mult = 1
            if rest and rest[0].isdigit():
                i = 0
                while i < len(rest) and rest[i].isdigit():
                    i += 1
                mult = int(rest[:i])
                rest = rest[i:]
            for _ in range(mult):
                rec(group, level, parent)
            if rest:
                rec(rest, level, parent)
        else:
            raise SetupError(f"Not a recognized character: {ch}")

    rec(sequence, 1, 0)
    n = len(monomers)
    # convert to 0-based and trim the trailing extra entries
    ch0 = [[c - 1 for c in children[i]] for i in range(n)]
    pa0 = [[p - 1 for p in parents[i]] for i in range(n)]
    lv0 = [[s - 1 for s in lvl] for lvl in levels]
    return Topology("".join(monomers), ch0, pa0, lv0)


def max_distance_matrix(sequence: Sequence[int], topo: Topology, diameters: np.ndarray) -> np.ndarray:
    """BFS path lengths (sum of mean diameters along bonds); freely_jointed.jl:34-114."""
    n = len(sequence)
    out = np.zeros((n, n))
    for start in range(n):
        dist = np.zeros(n)
        queue = [(start, 0.0)]
        visited = {start}
        while queue:
            cur, d = queue.pop(0)
            for nb in topo.children[cur] + topo.parents[cur]:
                if nb not in visited and nb < n:
                    dij = (diameters[sequence[cur]] + diameters[sequence[nb]]) / 2.0
                    nd = d + dij
                    dist[nb] = max(dist[nb], nd)
                    queue.append((nb, nd))
                    visited.add(nb)
        out[start] = dist
    np.fill_diagonal(out, 0.0)
    return out


# ---------------------------------------------------------------- chemistry

def _state_families(entries: List[dict], roles: Dict[str, str], reactions: List[Reaction]) -> Dict[str, List[str]]:
    """build_state_families (state_families.jl:1-90)."""
    # neutralizers: species in reactions with an empty side
    for rxn in reactions:
        reac = [s for s, c in zip(rxn.species, rxn.coeffs) if c == -1]
        prod = [s for s, c in zip(rxn.species, rxn.coeffs) if c == +1]
        if not reac or not prod:
            for s in (prod if not reac else reac):
                if s in roles:
                    roles[s] = "neutralizer"
    polymer_monomers = {c for e in entries if len(e["sequence"]) > 1 for c in e["sequence"] if c.isalpha()}
    neighbors: Dict[str, set] = {}
    pending = []
    for rxn in reactions:
        reac = [s for s, c in zip(rxn.species, rxn.coeffs) if c == -1]
        prod = [s for s, c in zip(rxn.species, rxn.coeffs) if c == +1]
        if not reac or not prod:
            for s in (prod if not reac else reac):
                if s in polymer_monomers:
                    raise SetupError(f"Monomer '{s}' appears in a neutralization reaction but is part of a polymer")
            continue
        r, p = len(reac), len(prod)
        if not ((r == 1 and p in (1, 2)) or (r == 2 and p == 1)):
            raise SetupError(f"Unsupported reaction form: {rxn.species}")
        if r == 2 and p == 1 and all(roles.get(s, "monomer") == "monomer" for s in reac):
            raise SetupError(f"Reaction {rxn.species}: both reactants are monomers")
        if r == 1 and p == 2 and all(roles.get(s, "monomer") == "monomer" for s in prod):
            raise SetupError(f"Reaction {rxn.species}: both products are monomers")
        lhs = [s for s in reac if roles.get(s, "monomer") not in ("partner", "neutralizer")]
        rhs = [s for s in prod if roles.get(s, "monomer") not in ("partner", "neutralizer")]
        tot = len(lhs) + len(rhs)
        if tot > 2:
            raise SetupError(f"Reaction {rxn.species} connects more than two monomers")
        if tot == 2:
            a, b = lhs[0], rhs[0]
            neighbors.setdefault(a, set()).add(b)
            neighbors.setdefault(b, set()).add(a)
        elif tot == 1:
            pending.append((reac, prod))
    for reac, prod in pending:
        cands = [s for s in dict.fromkeys(reac + prod) if roles.get(s) == "monomer"]
        if len(cands) != 2:
            continue
        a, b = cands
        neighbors.setdefault(a, set())
        neighbors.setdefault(b, set())

        def comp(x):
            seen, st = set(), [x]
            while st:
                y = st.pop()
                if y not in seen:
                    seen.add(y)
                    st.extend(neighbors.get(y, ()))
            return seen
        if a in comp(b) or b in comp(a):
            continue
        neighbors[a].add(b)
        neighbors[b].add(a)
    all_monomers = {c for e in entries for c in e["sequence"] if c.isalpha()}
    visited: set = set()
    fams: Dict[str, List[str]] = {}
    for mon in sorted(all_monomers):
        if mon in visited:
            continue
        stack, group = [mon], []
        while stack:
            node = stack.pop()
            if node not in visited:
                visited.add(node)
                group.append(node)
                stack.extend(neighbors.get(node, ()))
        group = sorted(group)
        for m in group:
            fams[m] = group
    return fams


def _delta_muH(monomers: List[str], families: Dict[str, List[str]], reactions: List[Reaction],
               names: List[str], densities: List[float], pSpecies: List[float],
               valences: np.ndarray, constants: Constants) -> np.ndarray:
    """assign_delta_muH! (assign_delta_muH.jl:50-247); mutates pSpecies and reaction coeffs like the reference."""
    idx_of = {m: i for i, m in enumerate(monomers)}
    name_to_index = {n: i for i, n in enumerate(names)}
    dmu = np.zeros(len(monomers))
    # Step 1: neutralization pSpecies completion
    for rxn in reactions:
        reac = [s for s, c in zip(rxn.species, rxn.coeffs) if c == -1]
        prod = [s for s, c in zip(rxn.species, rxn.coeffs) if c == +1]
        if not reac or not prod:
            m1, m2 = (prod if not reac else reac)
            tot = -LOG10 * rxn.pK
            i1, i2 = name_to_index[m1], name_to_index[m2]
            p1, p2 = pSpecies[i1], pSpecies[i2]
            if not math.isnan(p1):
                pSpecies[i2] = (tot - (-LOG10 * p1)) / -LOG10
            elif not math.isnan(p2):
                pSpecies[i1] = (tot - (-LOG10 * p2)) / -LOG10
    # Step 2: graph of delta mu between family members
    graph: Dict[str, List[Tuple[str, float]]] = {}
    for rxn in reactions:
        reac = [s for s, c in zip(rxn.species, rxn.coeffs) if c == -1]
        prod = [s for s, c in zip(rxn.species, rxn.coeffs) if c == +1]
        dq = sum(c * valences[idx_of[s]] for s, c in zip(rxn.species, rxn.coeffs) if s in idx_of)
        if abs(dq) > 1e-6:
            raise SetupError(f"Reaction {rxn.species} violates charge neutrality")
        sr = [s for s in reac if s in families]
        sp = [s for s in prod if s in families]
        found, m1, m2, fam1 = False, "", "", []
        for r in sr:
            for p in sp:
                m1, m2 = r, p
                fam1, fam2 = families[m1], families[m2]
                if fam1 == fam2:
                    found = True
                    break
            if found:
                break
        if not found:
            continue
        base = sorted(fam1)[0]
        fwd = (m1 == base)
        frm, to = (m1, m2) if fwd else (m2, m1)
        d = LOG10 * rxn.pK
        for s, c in zip(rxn.species, rxn.coeffs):
            if s == m1 or s == m2:
                continue
            if s not in name_to_index:
                raise SetupError(f"Context species {s} is not defined in species list.")
            ps = pSpecies[name_to_index[s]]
            if not math.isnan(ps):
                d -= LOG10 * ps * c
            else:
                rho = densities[name_to_index[s]]
                if rho > 0:
                    d += c * -math.log(constants.density_to_concentration(rho))
                else:
                    d = 0.0
        graph.setdefault(frm, []).append((to, d))
        graph.setdefault(to, []).append((frm, d))
        rxn.coeffs = [0] * len(rxn.coeffs)
    done = set()
    for fam in families.values():
        key = tuple(fam)
        if key in done:
            continue
        done.add(key)
        base = sorted(fam)[0]
        work = {base: 0.0}
        visited = set()

        def dfs(m, acc):
            visited.add(m)
            for nb, dm in graph.get(m, []):
                if nb not in visited:
                    work[nb] = acc + dm
                    dfs(nb, acc + dm)
        dfs(base, 0.0)
        for m, v in work.items():
            dmu[idx_of[m]] = v
        dmu[idx_of[base]] = 0.0
    # final pass: drop neutralization reactions whose pSpecies are known
    for rxn in reactions:
        reac = [s for s, c in zip(rxn.species, rxn.coeffs) if c == -1]
        prod = [s for s, c in zip(rxn.species, rxn.coeffs) if c == +1]
        if (not reac or not prod) and all(c != 0 for c in rxn.coeffs):
            m1, m2 = (prod if not reac else reac)
            p1, p2 = pSpecies[name_to_index[m1]], pSpecies[name_to_index[m2]]
            if not math.isnan(p1) or not math.isnan(p2):
                rxn.coeffs = [0] * len(rxn.coeffs)
    return dmu


def build_molsys(dataset: dict) -> MolSys:
    """setup_molsys (src/molecular_system/setup/setup_molsys.jl:48-75)."""
    ru = dataset.get("reduced_units") or {}
    constants = Constants(l_sc=float(ru["l_sc"]) * 1e-9 if "l_sc" in ru else 0.5e-9)
    # Temperature and dielectric are read from `constraints:` only
    # (get_temperature.jl:11-23, get_dielectric_constant.jl:11-25).
    cons = dataset.get("constraints") or {}
    T = float(cons.get("temperature", 298.15))
    er = float(cons.get("dielectric_constant", 78.4))
    lB = E0 ** 2 / (4 * math.pi * EPSILON0 * er * K_B * T) / constants.l_sc

    models = dataset["model"]
    models = [models] if isinstance(models, str) else list(models)
    # `coul` is listed by every shipped input but registered nowhere; Poisson
    # is always solved by the SCF path, so it is accepted and ignored here.
    fe_model = [m for m in models if m != "coul"]
    for m in fe_model:
        if m not in KNOWN_FE_MODELS:
            raise SetupError(f"Unknown model key: {m}")

    reactions = []
    for r in dataset.get("reactions") or []:
        coeffs = []
        for c in r["coeffs"]:
            ci = int(round(float(c)))
            if abs(float(c) - ci) > 1e-8 or ci not in (1, -1):
                raise SetupError(f"Unsupported coefficient {c}")
            coeffs.append(ci)
        if len(coeffs) != len(r["species"]):
            raise SetupError("Mismatch between species and coefficients")
        reactions.append(Reaction(list(r["species"]), coeffs, float(r["pK"])))

    entries = dataset["species"]
    names, dens, pS, adj, roles = [], [], [], [], {}
    for e in entries:
        name = e["sequence"]
        has_d, has_c, has_p = "density" in e, "concentration" in e, "pSpecies" in e
        if (has_d or has_c) and has_p:
            raise SetupError(f"Species '{name}' is overspecified")
        if has_d and has_c:
            raise SetupError(f"Species '{name}' has both density and concentration")
        if has_p and e.get("adjustable", False):
            raise SetupError(f"Species '{name}' has pSpecies and adjustable")
        names.append(name)
        if has_c:
            dens.append(constants.concentration_to_density(float(e["concentration"])))
        else:
            dens.append(float(e.get("density", 0.0)))
        pS.append(float(e["pSpecies"]) if has_p else float("nan"))
        adj.append(bool(e.get("adjustable", False)))
        roles[name] = str(e.get("role", "monomer"))

    used = set(c for e in entries for c in e["sequence"] if c.isalpha())
    for r in reactions:
        used.update(r.species)
    monomers = sorted(used)  # Python str sort == Julia codepoint sort
    idx_of = {m: i for i, m in enumerate(monomers)}

    families = _state_families(entries, roles, reactions)

    mdata = dataset["monomers"]
    diam = np.zeros(len(monomers))
    val = np.zeros(len(monomers))
    for m in monomers:
        if m not in mdata:
            raise SetupError(f"Species {m} listed in used_monomers not found in monomers")
        diam[idx_of[m]] = float(mdata[m].get("diameter", 0.0))
        val[idx_of[m]] = float(mdata[m].get("valence", 0.0))

    species: List[Species] = []
    for k, e in enumerate(entries):
        topo = parse_topology(e["sequence"])
        seq = [idx_of[c] for c in topo.monomers]
        fam = []
        for c in topo.monomers:
            if c in families:
                fam.append([idx_of[n] for n in families[c]])
            else:
                fam.append([idx_of[c]])
        model = e.get("model", "freely_jointed")
        if model not in ("freely_jointed", "monomerbead"):
            raise SetupError(f"Unknown chain model {model}")
        species.append(Species(
            name=e["sequence"], sequence=seq, state_family=fam, topology=topo, model=model,
            max_distance=max_distance_matrix(seq, topo, diam) if model == "freely_jointed" else None,
            input_density=dens[k], pSpecies=pS[k], adjustable=adj[k], role=roles[e["sequence"]],
            fixed=e.get("fixed"), evaluation=e.get("evaluation", "analytical")))

    bonds = set()
    for sp in species:
        for lvl in sp.topology.levels:
            for child in lvl:
                for j in sp.state_family[child]:
                    for parent in sp.topology.parents[child]:
                        for j1 in sp.state_family[parent]:
                            bonds.add((min(j, j1), max(j, j1)))
    bond_types = sorted(bonds)

    dmu = _delta_muH(monomers, families, reactions, names, dens, pS, val, constants)
    for k, sp in enumerate(species):
        sp.pSpecies = pS[k]

    sw = process_square_well(monomers, dataset)
    if "sw" in fe_model and (sw is None or not sw.pairs):
        raise SetupError("model 'sw' needs pairs: square_well entries")
    return MolSys(constants=constants, temperature=T, dielectric=er, bjerrum=lB,
                  monomers=monomers, diameters=diam, valences=val, delta_muH=dmu,
                  species=species, reactions=reactions, bond_types=bond_types, fe_model=fe_model,
                  sw=sw if "sw" in fe_model else None)
