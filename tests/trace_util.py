"""Readers of the reference's trace files, following tools/ttor_logging.py:11-31 (make_DAG_nodes,
make_DAG_edges) and :94-108 (make_trace): a Logger CSV "what,start,end" whose `what` is
"<who>><event>><task>", and a DepsLogger CSV "src,dst"."""
import pandas as pd


def dag_nodes(path):
    """Task rows as make_DAG_nodes builds them: '>run>' events, 'intern' ones dropped, task = 3rd '>' field."""
    df = pd.read_csv(path, header=None, names=["what", "time_start", "time_end"])
    df = df[df["what"].str.contains(">run>")]
    df = df[~df["what"].str.contains("intern")]
    cols = df["what"].str.split(pat=">", expand=True)
    df = df.assign(task=cols[2], who=cols[0])
    df = df.assign(kind=[s.split("_")[0] for s in df["task"]])
    return df


def dag_edges(path):
    return pd.read_csv(path, header=None, names=["start", "end"])


def trace_rows(path):
    """make_trace's split of every row into who / what / details."""
    df = pd.read_csv(path, sep=",", names=["info", "start", "end"], header=None)
    ss = [row.split(">") for row in df["info"]]
    return df.assign(who=[r[0] for r in ss], what=[r[1] for r in ss], details=[r[2:] for r in ss])
