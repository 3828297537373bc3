#!/bin/bash
# Usage: ./process-batch-query.sh <graph> <gpu> <pathToQueryDir>   (reference: system/process-batch-query.sh)
if [ "$#" -ne 3 ]; then
    echo "Usage: ./process-batch-query.sh <graph> <gpu> <pathToQueryDir>"
    exit 2
fi
cd "$(dirname "$0")/../.." && python3 -m everest_amd.system.batch_query "$1" "$2" "$3" | python3 -m everest_amd.system.filter
