'''

Writes host_freeze.json: SHA-256 digests of what the host-side scenario code produces, so that a
rewrite of those modules (particles, bathymetry, boundaries, refinement, breaking) can be held to
its outputs bit for bit. Every golden file of the breaking tests and every oracle comparison starts
from these states, so they must not move.

Covered: the starts of every configuration (first-order and exact solitary waves, the 64 ppd single
run, sweep runs, 3D ridges with and without a split crest zone, a serving ridge) with their boundary
descriptions; the beds' elevation, depth, slope, normal and samples; the height fields' elevation,
gradient, normal and sections; the boundaries' layers, columns and height-field clamp; jittered
lattices; a split and a merge; and the breaking read-out's measures and a tracker's whole history
over a wave steepened and overturned by hand.

Run from the repository root: python tests/golden/make_host_freeze.py (--check compares instead).

'''

import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'host_freeze.json')


def digest(value) -> str:
    '''A SHA-256 of a value's exact contents: arrays by dtype, shape and bytes, floats by their hex.'''
    h = hashlib.sha256()

    def feed(v):
        if v is None:
            h.update(b'N')
        elif isinstance(v, dict):
            h.update(b'D')
            for key in sorted(v):
                h.update(str(key).encode())
                feed(v[key])
        elif isinstance(v, (list, tuple)):
            h.update(b'L%d' % len(v))
            for item in v:
                feed(item)
        elif isinstance(v, np.ndarray):
            a = np.ascontiguousarray(v)
            h.update(f'A{a.dtype.str}{a.shape}'.encode())
            h.update(a.tobytes())
        elif isinstance(v, (bool, np.bool_)):
            h.update(b'T' if v else b'F')
        elif isinstance(v, (int, np.integer)):
            h.update(b'I%d' % int(v))
        elif isinstance(v, (float, np.floating)):
            h.update(b'F' + float(v).hex().encode())
        elif isinstance(v, str):
            h.update(b'S' + v.encode())
        elif hasattr(v, '__dict__'):
            # a record of the read-out (a surface, a face): its fields
            h.update(b'O' + type(v).__name__.encode())
            feed(dict(vars(v)))
        else:
            raise TypeError(f'cannot digest {type(v)}')

    feed(value)
    return h.hexdigest()


def caseState(case) -> dict:
    p = case.particles
    return {'positions': p.positions, 'velocities': p.velocities, 'densities': p.densities, 'masses': p.masses,
            'isFluid': p.isFluid, 'lengths': p.lengths, 'boundary': case.boundary.description(),
            'h': case.smoothingLength, 'domain': [case.domainMin, case.domainMax], 'c0': case.soundSpeed}


def scenarios() -> dict:
    from pysurf_amd import scenarios as sc
    out = {}
    out['solitary16'] = caseState(sc.solitaryOnSlope(particlesPerDepth = 16))
    out['solitary8seed1'] = caseState(sc.solitaryOnSlope(particlesPerDepth = 8, seed = 1, waveHeight = 0.3))
    out['single64'] = caseState(sc.singleRun(seed = 0, particlesPerDepth = 64))
    for seed in range(6):
        out[f'sweep{seed}'] = caseState(sc.sweepRun(seed))
    out['ridgeSplit'] = caseState(sc.ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5,
                                             ridgeWidth = 0.6, zoneHalfWidth = 0.4, lengthRatio = 0.5, seed = 0))
    out['ridgeWhole'] = caseState(sc.ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5,
                                             ridgeWidth = 0.6, zoneHalfWidth = None))
    out['ridgeKeep'] = caseState(sc.ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5,
                                            ridgeWidth = 0.6, zoneHalfWidth = 0.4, lengthRatio = 1.0))
    out['servingRidge0'] = caseState(sc.servingRidge(0))
    out['grilli64'] = caseState(sc.exactSolitaryOnSlope(64))
    out['exact32jitter'] = caseState(sc.exactSolitaryOnSlope(32, jitter = 0.1, seed = 3, bedFitted = False))
    out['exactRidge'] = caseState(sc.ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5,
                                             ridgeWidth = 0.6, zoneHalfWidth = None, start = 'exact'))
    return out


def beds() -> dict:
    from pysurf_amd.bathymetry import Bathymetry, HeightField
    out = {}
    x = np.linspace(-1.0, 31.0, 997)
    for name, bed in (('plane', Bathymetry.planeBeach(depth = 1.0, toeX = 8.0, slope = 1.0 / 15.0, endX = 25.0)),
                      ('steep', Bathymetry.planeBeach(depth = 0.7, toeX = 3.0, slope = 1.0 / 8.0, endX = 12.0, startX = -2.0))):
        out[name] = {'elevation': bed.elevation(x), 'depth': bed.depth(x), 'slope': bed.slope(x), 'normal': bed.normal(x),
                     'sample': bed.sample(0.05), 'startX': bed.startX, 'endX': bed.endX}
    rng = np.random.default_rng(5)
    px, py = rng.uniform(-1.0, 21.0, 500), rng.uniform(-6.0, 6.0, 500)
    fields = {'gaussian': HeightField.gaussianRidge(depth = 1.0, height = 0.5, width = 2.0, slope = 1.0 / 15.0, toeX = 6.0,
                                                    length = 20.0, halfWidth = 5.0, direction = 25.0),
              'sech': HeightField.sechRidge(0.8)}
    for name, field in fields.items():
        out[name] = {'elevation': field.elevation(px, py), 'gradient': field.gradient(px, py), 'normal': field.normal(px, py),
                     'depth': field.depth(px, py), 'lowest': field.lowestElevation(), 'bounds': list(field.bounds),
                     'section': field.section(0.3).sample(0.05), 'table': field.z}
    return out


def boundaries() -> dict:
    from pysurf_amd.bathymetry import Bathymetry, HeightField
    from pysurf_amd.boundaries import BedBoundary, ContainerBoundary, HeightFieldBoundary
    from pysurf_amd.particles import ParticleSystem
    out = {}
    out['container2'] = ContainerBoundary([0.0, 0.0], [1.0, 0.6], 0.05, nLayers = 3).positions()
    out['container3'] = ContainerBoundary([0.0, 0.0, 0.0], [0.5, 0.3, 0.4], 0.05, nLayers = 3, openTop = False).positions()
    bed = Bathymetry.planeBeach(depth = 1.0, toeX = 4.0, slope = 1.0 / 10.0, endX = 12.0)
    for name, b in (('bed', BedBoundary(bed, 0.05, nLayers = 3, wallTop = 1.5, endWallTop = 1.5)),
                    ('bedOpen', BedBoundary(bed, 0.04, nLayers = 4))):
        out[name] = {'positions': b.positions(), 'description': b.description()}
    field = HeightField.gaussianRidge(depth = 1.0, height = 0.5, width = 0.6, slope = 1.0 / 15.0, toeX = 2.0, length = 6.0,
                                      halfWidth = 0.75, direction = 25.0)
    hb = HeightFieldBoundary(field, 0.1, nLayers = 3, wallTop = 1.6)
    rng = np.random.default_rng(7)
    cloud = np.column_stack([rng.uniform(-0.2, 6.2, 400), rng.uniform(-0.9, 0.9, 400), rng.uniform(-0.2, 1.2, 400)])
    particles = ParticleSystem.assemble(cloud, np.zeros((0, 3)), 0.1, 1000.0, rng.normal(size = (400, 3)))
    hb.enforce(particles)
    out['heightField'] = {'positions': hb.positions(), 'columns': hb.columns(), 'floor': hb.latticeFloor,
                          'description': hb.description(), 'enforced': [particles.positions, particles.velocities]}
    return out


def jitters() -> dict:
    from pysurf_amd.particles import jitterPoints, latticePoints
    out = {}
    points = latticePoints([0.0, 0.0], [2.0, 1.0], 0.05)
    floor = lambda x: 0.1 * np.sin(3.0 * x)
    out['2d'] = jitterPoints(points, 0.05, 0.3, 4, [0.0], [2.0], floor, lambda x: 0.95 + 0.02 * x)
    out['2dNoCeiling'] = jitterPoints(points, 0.05, 0.45, 5, [0.0], [2.0], floor)
    points3 = latticePoints([0.0, 0.0, 0.0], [0.6, 0.4, 0.5], 0.05)
    out['3d'] = jitterPoints(points3, 0.05, 0.4, 6, [0.0, 0.0], [0.6, 0.4], lambda x, y: 0.05 * x + 0.02 * y,
                             lambda x, y: 0.45 + 0.0 * x)
    return out


def refinements() -> dict:
    from pysurf_amd import scenarios as sc
    from pysurf_amd.refinement import mergeParticles, refinementLevels, splitParticles
    out = {}
    case = sc.solitaryOnSlope(particlesPerDepth = 8, seed = 2)
    p = case.particles.copy()
    unsplit = float(p.masses[p.isFluid][0])
    x = p.positions[:, 0]
    chosen = splitParticles(p, p.isFluid & (x > 3.0) & (x < 5.0), case.referenceDensity, case.smoothingLength, 0.5)
    out['split'] = {'chosen': chosen, 'positions': p.positions, 'masses': p.masses, 'lengths': p.lengths,
                    'levels': refinementLevels(p, unsplit)}
    rng = np.random.default_rng(8)
    p.positions = p.positions + rng.uniform(-0.01, 0.01, p.positions.shape) * case.spacing
    p.velocities = p.velocities + rng.normal(scale = 0.05, size = p.velocities.shape)
    x = p.positions[:, 0]
    result = mergeParticles(p, p.isFluid & (x > 3.0) & (x < 5.0), case.referenceDensity, unsplit, case.smoothingLength, 0.5)
    out['merge'] = {'groups': result.groups, 'energy': result.removedEnergy, 'source': result.source,
                    'positions': p.positions, 'velocities': p.velocities, 'masses': p.masses, 'densities': p.densities,
                    'lengths': p.lengths, 'isFluid': p.isFluid}
    ridge = sc.ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                       zoneHalfWidth = None)
    q = ridge.particles.copy()
    unsplit3 = float(q.masses[q.isFluid][0])
    x = q.positions[:, 0]
    splitParticles(q, q.isFluid & (x > 2.0) & (x < 3.0), ridge.referenceDensity, ridge.smoothingLength, 0.5)
    q.positions = q.positions + rng.uniform(-0.01, 0.01, q.positions.shape) * ridge.spacing
    x = q.positions[:, 0]
    result = mergeParticles(q, q.isFluid & (x > 2.0) & (x < 3.0), ridge.referenceDensity, unsplit3, ridge.smoothingLength,
                            0.5)
    out['merge3'] = {'groups': result.groups, 'energy': result.removedEnergy, 'positions': q.positions, 'masses': q.masses}
    return out


def wave(time: float, spacing: float = 1.0 / 16.0) -> np.ndarray:
    '''Water over a 1:15 beach whose crest travels shoreward, steepens and at last leans over its front: by hand.'''
    x = np.arange(spacing / 2.0, 24.0, spacing)
    columns = []
    crest = 4.0 + 3.6 * time
    height = 0.45 + 0.05 * time
    for xi in x:
        bed = 0.0 if xi < 8.0 else (xi - 8.0) / 15.0
        lean = 0.25 * max(time - 2.5, 0.0)
        d = xi - crest + lean * (1.0 if xi > crest else 0.0)
        surface = 1.0 + height / np.cosh(1.2 * d * (1.0 + 0.4 * time))**2
        z = np.arange(bed + spacing / 2.0, surface, spacing)
        columns.append(np.column_stack([np.full(len(z), xi), z]))
    points = np.vstack(columns)
    # the overturning lip: a tongue of particles thrown ahead of the crest after t = 3
    if time > 3.0:
        tip = np.arange(crest, crest + 0.6 * (time - 3.0), spacing)
        tongue = np.column_stack([tip, 1.0 + height - 0.3 * (tip - crest)**2])
        points = np.vstack([points, tongue])
    return points


def breakingReadout() -> dict:
    from pysurf_amd import breaking as br
    from pysurf_amd.bathymetry import Bathymetry
    out = {}
    bed = Bathymetry.planeBeach(depth = 1.0, toeX = 8.0, slope = 1.0 / 15.0, endX = 25.0)
    s = 1.0 / 16.0
    positions = wave(3.4)
    out['profile'] = br.freeSurfaceProfile(positions, 0.0, 24.0, s)
    centers, elevation = out['profile']
    crestX = float(centers[np.nanargmax(elevation)])
    out['trough'] = br.troughAhead(centers, elevation, crestX, 3.0)
    out['lean'] = br.frontFaceLean(positions, crestX, 1.0, s, 1.0)
    out['cells'] = br.faceCells(positions, crestX, 1.0, s, 1.0)
    out['waterline'] = br.waterlineX(positions, bed, 0.05)
    grid = br.fractionGrid(s, [crestX - 1.5, 0.3], [crestX + 1.5, 1.8])
    out['grid'] = grid
    out['fraction'] = br.fluidFraction(positions, s, [crestX - 1.5, 0.3], [crestX + 1.5, 1.8])
    surface = br.particleSurface(positions, s, [crestX - 1.5, 0.3], [crestX + 1.5, 1.8])
    out['surface'] = {'surface': surface.surface, 'pockets': list(surface.pockets), 'drops': list(surface.drops)}
    face = br.contourFace(positions, s, crestX, bed, 1.0, 0.8)
    out['face'] = {key: getattr(face, key) for key in face.__dataclass_fields__}
    for measure in ('cells', 'contour'):
        tracker = br.BreakingTracker(bed, s, faceReach = 1.0, searchFromX = 8.0, faceMeasure = measure)
        for k in range(90):
            t = 0.05 * k
            tracker.update(t, wave(t))
            if tracker.breakingFinal:
                break
        history = {name: np.asarray(getattr(tracker, name), dtype = float) for name in
                   ('times', 'crestX', 'crestHeight', 'lean', 'faceSlope', 'waterline')}
        point = tracker.breakingPoint
        history['point'] = None if point is None else {key: getattr(point, key) for key in point.__dataclass_fields__}
        out[f'tracker-{measure}'] = history
    return out


def everything() -> dict:
    groups = {'scenarios': scenarios, 'beds': beds, 'boundaries': boundaries, 'jitters': jitters,
              'refinements': refinements, 'breaking': breakingReadout}
    record = {}
    for group, make in groups.items():
        values = make()
        for key, value in values.items():
            record[f'{group}/{key}'] = digest(value)
    return record


def main(check: bool) -> int:
    record = everything()
    if check:
        frozen = json.load(open(OUT))
        moved = [key for key in frozen if frozen[key] != record.get(key)]
        print('moved:', moved if moved else 'none')
        return 1 if moved else 0
    with open(OUT, 'w') as f:
        json.dump(record, f, indent = 1, sort_keys = True)
    print(f'{len(record)} digests written')
    return 0


if __name__ == '__main__':
    sys.exit(main('--check' in sys.argv))
