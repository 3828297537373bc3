#!/bin/bash
# Usage: ./process-single-query.sh <pathToQuery>   (reference: system/process-single-query.sh)
if [ "$#" -ne 1 ]; then
    echo "Usage: ./process-single-query.sh <pathToQuery>"
    exit 2
fi
cd "$(dirname "$0")/../.." && python3 -m everest_amd.system.single_query "$1" | python3 -m everest_amd.system.filter
