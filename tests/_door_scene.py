"""The door scene shared by the oracle pins (CPU) and the GPU door tests:
one door plate at x = 2 whose 19-bin table is penetrable (< 0) below 47.5
degrees of incidence and reflecting above, a Tx at (0, 0, 1) and Rx on both
sides; door_expected() enumerates its records and H from the geometry."""
import math

import numpy as np

WL = 3e8 / 60e9


def quad(p0, p1, p2, p3):
    return [list(p0) + list(p1) + list(p2), list(p0) + list(p2) + list(p3)]


def bin_of(theta):
    """nearest 5-degree bin, from the angle in radians (the bin centre
    k * 5 deg closest to theta)"""
    centres = np.radians(5.0 * np.arange(19))
    return int(np.argmin(np.abs(centres - theta)))


DOOR_X = 2.0
DOOR = quad((DOOR_X, -3.0, -1.0), (DOOR_X, 3.0, -1.0), (DOOR_X, 3.0, 3.0), (DOOR_X, -3.0, 3.0))
DOOR_TABLE = np.where(np.arange(19) < 10, -1.0, 1.0) * (0.5 + 0.02 * np.arange(19))
TX = np.array([0.0, 0.0, 1.0])
# far side (x = 4): incidence atan(|y| / 4) below 47.5 deg for |y| < 4.365,
# the door's edge at |y| = 6
FAR_Y = [0.0, 1.5, -2.2, 3.0, 4.0, -4.8, 5.5, 6.8, -7.5]
# near side (x = 1): the door reflection's specular point at y' = 2 y / 3,
# incidence atan(|y| / 3)
NEAR_Y = [0.3, -1.2, 2.0, 3.3, -4.1]
RX_R = 0.1


def door_rx():
    pts = [(4.0, y, 1.0) for y in FAR_Y] + [(1.0, y, 1.0) for y in NEAR_Y]
    return np.asarray(pts, np.float32)[None, :, :]


def door_expected():
    """(h, b, planes, tile) records and H of every Rx, from the geometry."""
    keys, H = set(), []
    rx = door_rx()[0].astype(np.float64)
    for j, r in enumerate(rx):
        hj = 0j
        if r[0] > DOOR_X:
            d = np.linalg.norm(r - TX)
            yc = TX[1] + (r[1] - TX[1]) * (DOOR_X - TX[0]) / (r[0] - TX[0])
            th = math.atan2(abs(r[1] - TX[1]), r[0] - TX[0])
            if abs(yc) > 3.0:  # past the door's edge: line of sight
                keys.add((0, 0, (), j))
                hj += WL / (4 * math.pi * d) * np.exp(-1j * 2 * math.pi * d / WL)
            elif DOOR_TABLE[bin_of(th)] < 0:  # transmitted through the door
                keys.add((1, 0, (), j))
                g = abs(float(np.float32(DOOR_TABLE[bin_of(th)])))
                hj += g * WL / (4 * math.pi * d) * np.exp(-1j * 2 * math.pi * d / WL)
            # else: the door reflects that ray back; nothing reaches this Rx
        else:
            d = np.linalg.norm(r - TX)
            keys.add((0, 0, (), j))
            hj += WL / (4 * math.pi * d) * np.exp(-1j * 2 * math.pi * d / WL)
            I = TX.copy()
            I[0] = 2 * DOOR_X - TX[0]
            L = np.linalg.norm(r - I)
            v = I - r
            th = math.acos(abs(v[0]) / np.linalg.norm(v))
            keys.add((1, 1, (0,), j))
            g = abs(float(np.float32(DOOR_TABLE[bin_of(th)])))
            hj += g * WL / (4 * math.pi * L) * np.exp(-1j * 2 * math.pi * L / WL)
        H.append(hj)
    return keys, np.asarray(H)
