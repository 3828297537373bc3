"""Keep only the lines between "@@@ [output starts]" and "@@@ [output ends]"
(reference: system/filter.py:1-15)."""
import sys


def main(stream=sys.stdin, out=sys.stdout):
    started = False
    for line in stream:
        line = line.strip()
        if line == "@@@ [output starts]":
            started = True
            continue
        if started:
            if line == "@@@ [output ends]":
                started = False
            else:
                print(line, file=out)


if __name__ == "__main__":
    main()
