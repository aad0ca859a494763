#!/bin/bash
# The cubic-code simulation pipeline of SOGRAND_C/cubic_flow.sh with the MI355X
# decoder: the same five steps, file names and formats; the decoder step runs
# the drop-in sogrand_amd/cubic_decoder. Files are written to the current
# directory. Settings: SNR_DB (default 2.0), N_CW codewords (default 1000),
# SEED (default 1), DEC_FLAGS (decoder options, e.g. "--variant legacy").
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
export PYTHONPATH="$ROOT${PYTHONPATH:+:$PYTHONPATH}"

ORIG_DATA="original_data.bin"
ENCODED_DATA="encoded_cubic.bin"
CORRUPTED_LLRS="corrupted_llrs_cubic.bin"
DECODED_DATA="decoded_cubic.bin"
SNR_DB="${SNR_DB:-2.0}"
N_CW="${N_CW:-1000}"
SEED="${SEED:-1}"

echo "--- Starting CUBIC Code Simulation Pipeline (SNR = ${SNR_DB} dB) ---"
if [ ! -x "$HERE/cubic_decoder" ]; then
    echo "Building the decoder..."
    python3 -c "import __graft_entry__ as g; g.build()"
fi
python3 -m sogrand_amd.flow data --n-cw "$N_CW" --seed "$SEED" "$ORIG_DATA"
echo "Encoding data..."
python3 -m sogrand_amd.flow encode "$ORIG_DATA" "$ENCODED_DATA"
echo "Simulating channel..."
python3 -m sogrand_amd.flow channel "$ENCODED_DATA" "$CORRUPTED_LLRS" "$SNR_DB" 3 --seed "$SEED"
echo "Decoding data..."
"$HERE/cubic_decoder" "$CORRUPTED_LLRS" "$DECODED_DATA" ${DEC_FLAGS:-}
echo "Comparing results..."
python3 -m sogrand_amd.flow compare "$ORIG_DATA" "$DECODED_DATA"
echo "--- Pipeline finished ---"
