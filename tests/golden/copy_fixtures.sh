#!/bin/sh
# Copies the small data files the tests use from the reference's examples
# (data only: inputs the reference ships for its own example runs).
set -e
REF=${REF:-/root/reference}
HERE=$(dirname "$0")
cp "$REF/doc/examples/LCLS/wakes/Sx_200um_6mm_xband.sdds" "$HERE/wake_Sx_200um_6mm_xband.sdds"
cp "$REF/doc/examples/PAR/tracking/parTrack.lte" "$HERE/parTrack.lte"
