#!/bin/sh
# Copies the reference's own input fixtures (data files its tests and run
# scripts use: data/1a07C1, data/edud, data/10gs, data/parameters) into
# tests/golden/inputs, with the distribution's licence (its condition 2 asks
# that recipients receive a copy; NOTICE.txt lists the files). Data only; no
# source text.
set -e
R=${1:-/root/reference/data}
D=$(dirname "$0")/inputs
mkdir -p "$D/1a07C1" "$D/edud" "$D/10gs" "$D/parameters"
cp "$R/1a07C1/1a07C1.sdf" "$R/1a07C1/1a07C.pdb" "$R/1a07C1/1a07C1-0.8.ff" "$D/1a07C1/"
cp "$R/edud/1b9v_4.sdf" "$R/edud/1e66_4.sdf" "$R/edud/1b9v_4.ff" "$D/edud/"
cp "$R/10gs/10gs_ligand.sdf" "$R/10gs/10gsA00.pdb" "$R/10gs/10gsA00.ff" "$D/10gs/"
cp "$R/parameters/paras" "$R/parameters/08ff_opt" "$R/parameters/08_nor_a" "$R/parameters/08_nor_b" "$R/parameters/trace.txt" "$D/parameters/"
cp "$R/trajectory.out" "$D/"
cp "$R/../LICENSE.txt" "$D/"
chmod u+w -R "$D"
