"""Reader of HepMC3 ASCII (Asciiv3) files for the tests: parses the header,
events, particles and vertices and rebuilds each event's vertex graph, the way
HepMC3's ReaderAscii resolves a P line's parent field (negative: an explicit
vertex from a V line; positive: the implicit end vertex of that particle;
0: no production vertex)."""
from dataclasses import dataclass, field


@dataclass
class Particle:
    id: int
    parent: int
    pdg: int
    p: tuple  # (px, py, pz, E)
    mass: float
    status: int


@dataclass
class Event:
    number: int
    n_vertices: int
    n_particles: int
    units: tuple = ()
    weights: list = field(default_factory=list)
    attributes: dict = field(default_factory=dict)
    particles: list = field(default_factory=list)
    vertices: dict = field(default_factory=dict)  # explicit: id -> (status, [incoming ids])

    def vertex_graph(self):
        """{vertex key: (incoming particle ids, outgoing particle ids)}. Explicit
        vertices are keyed by their negative id, implicit ones by ('end', pid)."""
        g = {vid: (list(inc), []) for vid, (_, inc) in self.vertices.items()}
        for q in self.particles:
            if q.parent < 0:
                g[q.parent][1].append(q.id)
            elif q.parent > 0:
                g.setdefault(("end", q.parent), ([q.parent], []))[1].append(q.id)
        return g


@dataclass
class HepMC3File:
    version: str
    tools: list
    events: list
    footer: bool


def unescape(s):
    out, i = [], 0
    while i < len(s):
        if s[i] == "\\" and i + 1 < len(s):
            out.append("\n" if s[i + 1] == "|" else s[i + 1])
            i += 2
        else:
            out.append(s[i])
            i += 1
    return "".join(out)


def read(path):
    with open(path) as fh:
        lines = fh.read().split("\n")
    if not lines[0].startswith("HepMC::Version "):
        raise ValueError("missing version line")
    if lines[1] != "HepMC::Asciiv3-START_EVENT_LISTING":
        raise ValueError("missing start line")
    version = lines[0].split()[1]
    tools, events, footer = [], [], False
    ev = None
    for ln in lines[2:]:
        if not ln:
            continue
        if ln == "HepMC::Asciiv3-END_EVENT_LISTING":
            footer = True
            break
        tag, rest = ln[0], ln[2:]
        if tag == "T":
            tools.append(unescape(rest).split("\n"))
        elif tag == "E":
            f = rest.split()
            ev = Event(int(f[0]), int(f[1]), int(f[2]))
            events.append(ev)
        elif tag == "U":
            ev.units = tuple(rest.split())
        elif tag == "W":
            ev.weights = [float(x) for x in rest.split()]
        elif tag == "A":
            f = rest.split(" ", 2)
            ev.attributes[f[1]] = unescape(f[2])
        elif tag == "V":
            f = rest.split()
            inc = f[2].strip("[]")
            ev.vertices[int(f[0])] = (int(f[1]), [int(x) for x in inc.split(",") if x])
        elif tag == "P":
            f = rest.split()
            ev.particles.append(Particle(int(f[0]), int(f[1]), int(f[2]),
                                         tuple(float(x) for x in f[3:7]), float(f[7]), int(f[8])))
        else:
            raise ValueError(f"unknown line: {ln[:40]}")
    return HepMC3File(version, tools, events, footer)
