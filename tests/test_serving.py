'''

The request loop's admission on the CPU: the fleets are replaced by stand-ins that advance each
admitted run one unit of time per reading, so the rules are checked without a device. A request
that can never fit the HBM budget is refused on arrival and the loop still returns; a request that
waits for room does not hold back a smaller one behind it; the queue depth is recorded.

'''

import pytest

from pysurf_amd.scenarios import solitaryOnSlope
from pysurf_amd.serving import Request, Server, latencyPercentiles


class StandInFleet:

    '''A fleet that takes no device: each reading advances its unfinished members by one unit of time.'''

    def __init__(self) -> None:
        self.members = []
        self.batch = None
        self.joined = []

    def rebuild(self, leaving: list, joining: list) -> None:
        self.members = [member for member in self.members if member not in leaving] + list(joining)
        self.joined += [member.key.key for member in joining]
        self.batch = object() if self.members else None

    def launch(self) -> None:
        pass

    def collect(self) -> dict:
        return {'members': [(index, member, member.time, member.steps) for index, member in enumerate(self.members)
                            if not member.done]}

    def track(self, reading: dict, wall: float) -> list:
        finished = []
        for _, member, _, _ in reading['members']:
            member.time += 1.0
            member.steps += 1
            if member.time >= member.endTime:
                member.done = True
                member.doneTime, member.doneSteps, member.finishedWall = member.time, member.steps, wall
                finished.append(member)
        return finished


def standInServer(budget: float) -> tuple:
    server = Server(budgetBytes = budget)
    fleets = server.fleets

    def fleet(dimensions: int):
        return fleets.setdefault(dimensions, StandInFleet())

    server.fleet = fleet
    return server, fleets


@pytest.mark.timeout(60)
def testARequestThatCanNeverFitIsRefusedAndTheLoopReturns():
    small, large = solitaryOnSlope(particlesPerDepth = 6), solitaryOnSlope(particlesPerDepth = 10)
    server, _ = standInServer(budget = 1.5 * Server().footprint(small))
    assert server.footprint(large) > server.budgetBytes
    served = server.serve([Request(key = 0, arrival = 0.0, case = large, endTime = 2.0),
                           Request(key = 1, arrival = 0.0, case = small, endTime = 2.0)])
    records = {record.key: record for record in served}
    assert records[0].refused == 'budget' and records[0].steps == 0
    assert records[1].refused == '' and records[1].steps == 2
    # the refused request is left out of the latency percentiles
    assert latencyPercentiles(served)['p50'] == pytest.approx(records[1].latency)


@pytest.mark.timeout(60)
def testASmallerRequestPassesOneWaitingForRoom():
    a, b, c = (solitaryOnSlope(particlesPerDepth = ppd, seed = seed) for ppd, seed in ((10, 1), (8, 2), (6, 3)))
    probe = Server()
    fa, fb, fc = probe.footprint(a), probe.footprint(b), probe.footprint(c)
    assert fa > fb > fc
    # a and c fit together, a and b do not; b fits alone
    server, fleets = standInServer(budget = fa + fc)
    served = server.serve([Request(key = 'a', arrival = 0.0, case = a, endTime = 3.0),
                           Request(key = 'b', arrival = 0.0, case = b, endTime = 2.0),
                           Request(key = 'c', arrival = 0.0, case = c, endTime = 1.0)])
    assert sorted(record.key for record in served) == ['a', 'b', 'c']
    assert fleets[2].joined[:2] == ['a', 'c'] and fleets[2].joined[2] == 'b'
    records = {record.key: record for record in served}
    assert records['b'].admitted > records['a'].admitted
    assert max(depth for _, depth in server.queueDepth) == 1
    assert all(record.refused == '' for record in served)


@pytest.mark.timeout(60)
def testFleetsBySpacingShareTheSlots():
    # With a fleet per 2D spacing, runs of two spacings join two fleets; the dimension's slots hold
    # the particles of both, so a third run that would overflow them waits for one to finish
    a, b, c = (solitaryOnSlope(particlesPerDepth = ppd, seed = seed) for ppd, seed in ((8, 1), (10, 2), (8, 3)))
    server, fleets = standInServer(budget = 1e12)
    server.fleetsBySpacing = True
    server.slots2D = a.count + b.count
    served = server.serve([Request(key = 'a', arrival = 0.0, case = a, endTime = 2.0),
                           Request(key = 'b', arrival = 0.0, case = b, endTime = 3.0),
                           Request(key = 'c', arrival = 0.0, case = c, endTime = 1.0)])
    assert sorted(record.key for record in served) == ['a', 'b', 'c']
    assert set(fleets) == {server.fleetKey(a), server.fleetKey(b)} and server.fleetKey(a) == server.fleetKey(c)
    assert fleets[server.fleetKey(a)].joined == ['a', 'c'] and fleets[server.fleetKey(b)].joined == ['b']
    records = {record.key: record for record in served}
    assert records['c'].admitted > records['a'].admitted
    assert max(depth for _, depth in server.queueDepth) == 1
